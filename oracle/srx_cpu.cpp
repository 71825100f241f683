// ORACLE -- test infrastructure only.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load libsrx_cpu.so, as the checker and as the timed CPU baseline; the product
// (salviarenderer_amd, libsrx_hip.so) never includes, links or calls anything here.
//
// CPU restatement of the Salvia software rasterizer's hot path (wuye9036/SalviaRenderer,
// salvia/src/core and salvia/src/resource), exported through the same C ABI as the product
// (include/srx.h).  It keeps the reference's structure so that it can stand in for it:
//  - std::thread pool of hardware_concurrency() threads claiming fixed-size packages
//    (eflib/include/eflib/concurrency/thread_context.h:16-102);
//  - per-thread 128-entry vertex caches (default_vertex_cache.cpp:334-411);
//  - clipping in packages of 8 with a sparse-then-compacted vertex list (geom_setup_engine.cpp);
//  - viewport transform of unique vertex pointers (rasterizer.cpp:1203-1243);
//  - 64x64 binning into per-thread tile lists, per-tile merge + sort (rasterizer.cpp:775-994);
//  - Larrabee 64->16->4->1 traversal with the SSE evaluation order (rasterizer.cpp:298-773);
//  - quad interpolation, early-Z, cpp pixel shaders and the framebuffer (framebuffer.cpp).
//
// Pinning (tests/test_oracle_pins.py, tests/test_golden.py): coverage by closed-form rectangles and
// a per-pixel brute-force evaluation of the decision chain; depth by the screen-space plane;
// RGBA8 rounding, resolve and the mip box filter by known answers; fast_log2/floori/ceili/round by
// their formulas; the EWA table by the reference's literals; thread-count invariance.
// Parity unpinned (no reference test, vector or build can check them here; they follow the cited
// SSE code line by line): the SSE addressers' cvtt + size*8192 float modulo (sampler.cpp:27-119),
// calc_anisotropic_info and the EWA probe loop (sampler.cpp:730-760, 875-958), calc_lod and the
// linear-mip blend (sampler.cpp:521-603), the Lambert shader's SASL derivative pairing
// (cgs_simd.cpp:267-312), compute_triangle_info's choice of v0 (rasterizer.cpp:876-899), and
// vs_invocations (cache- and schedule-dependent in the reference).
#include "../include/srx.h"

#include "oracle_sampler.h"
#include "oracle_sasl.h"
#include "../salviarenderer_amd/csrc/srx_sasl.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

using namespace srxo;

namespace srxo {

// Blender's EWA weight table (exp(a r) - exp(a)) / (1 - exp(a)), a = -2, r = i/255, as printed with
// six significant digits (the reference's table, sampler.cpp:622-677, holds exactly these values).
static float* make_ewa() {
  static float t[256];
  for (int i = 0; i < 256; ++i) {
    double r = i / 255.0;
    double w = (std::exp(-2.0 * r) - std::exp(-2.0)) / (1.0 - std::exp(-2.0));
    char buf[64];
    std::snprintf(buf, sizeof buf, "%g", w);
    t[i] = std::strtof(buf, nullptr);
  }
  return t;
}
static const float* ewa_table = make_ewa();
}  // namespace srxo

namespace {

thread_local int32_t g_last_error = 0;
thread_local std::string g_last_error_str;
int32_t fail(int32_t code, const char* msg) {
  g_last_error = 0;
  g_last_error_str = msg;
  return code;
}

// ---------------------------------------------------------------- thread pool
// execute_threads (thread_context.h:82-102): items split in packages claimed by an atomic cursor;
// the caller participates.
class pool {
public:
  explicit pool(unsigned n) : n_(n ? n : 1) {
    for (unsigned i = 1; i < n_; ++i) workers_.emplace_back([this, i] { loop(i); });
  }
  ~pool() {
    {
      std::lock_guard<std::mutex> lk(m_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : workers_) t.join();
  }
  unsigned size() const { return n_; }
  // fn(thread_id, package_begin, package_end)
  void run(size_t items, size_t package, const std::function<void(unsigned, size_t, size_t)>& fn) {
    if (items == 0) return;
    std::unique_lock<std::mutex> lk(run_m_);
    fn_ = &fn;
    items_ = items;
    package_ = package;
    cursor_.store(0);
    {
      std::lock_guard<std::mutex> lk2(m_);
      pending_ = n_ - 1;
      ++gen_;
    }
    cv_.notify_all();
    work(0);
    std::unique_lock<std::mutex> lk3(m_);
    done_cv_.wait(lk3, [this] { return pending_ == 0; });
  }

private:
  void work(unsigned tid) {
    for (;;) {
      size_t p = cursor_.fetch_add(1);
      size_t b = p * package_;
      if (b >= items_) break;
      size_t e = std::min(items_, b + package_);
      (*fn_)(tid, b, e);
    }
  }
  void loop(unsigned tid) {
    uint64_t seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(m_);
        cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
        if (stop_) return;
        seen = gen_;
      }
      work(tid);
      {
        std::lock_guard<std::mutex> lk(m_);
        if (--pending_ == 0) done_cv_.notify_all();
      }
    }
  }
  unsigned n_;
  std::vector<std::thread> workers_;
  std::mutex m_, run_m_;
  std::condition_variable cv_, done_cv_;
  bool stop_ = false;
  uint64_t gen_ = 0;
  unsigned pending_ = 0;
  const std::function<void(unsigned, size_t, size_t)>* fn_ = nullptr;
  size_t items_ = 0, package_ = 1;
  std::atomic<size_t> cursor_{0};
};

unsigned thread_count() {
  if (const char* e = std::getenv("SRX_CPU_THREADS")) {
    int n = std::atoi(e);
    if (n > 0) return static_cast<unsigned>(n);
  }
  unsigned n = std::thread::hardware_concurrency();
  if (n == 0) n = 1;
  // a host share smaller than the machine (OMP_NUM_THREADS) caps the pool
  if (const char* e = std::getenv("OMP_NUM_THREADS")) {
    int m = std::atoi(e);
    if (m > 0 && static_cast<unsigned>(m) < n) n = static_cast<unsigned>(m);
  }
  return n;
}

pool& global_pool() {
  static pool p(thread_count());
  return p;
}

// ---------------------------------------------------------------- objects
enum obj_kind { K_RENDERER = 1, K_BUFFER, K_TEXTURE, K_SURFACE, K_SAMPLER, K_LAYOUT, K_RASTER, K_DS, K_QUERY, K_BATCH, K_SERVER,
                K_SHADER, K_SWAP_CHAIN };

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

struct buffer_obj : object {
  std::vector<uint8_t> data;
  buffer_obj() : object(K_BUFFER) {}
};

struct surface_obj : object {
  int w, h, samples, pf;
  std::vector<uint8_t> data;
  surface_obj(int w_, int h_, int s_, int pf_)
      : object(K_SURFACE), w(w_), h(h_), samples(s_), pf(pf_),
        data(static_cast<size_t>(w_) * h_ * s_ * pf_size(pf_), 0) {}
  size_t esize() const { return pf_size(pf); }
  uint8_t* addr(size_t x, size_t y, size_t s) { return data.data() + ((y * w + x) * samples + s) * esize(); }
  vec4 get(size_t x, size_t y, size_t s) { return load_texel(pf, addr(x, y, s)); }
  void set(size_t x, size_t y, size_t s, vec4 c) { store_texel(pf, addr(x, y, s), c); }
  surf_view view() const { return surf_view{w, h, samples, pf, data.data()}; }
  // surface::fill (surface.cpp:170-275)
  void fill(vec4 c) {
    uint8_t px[16];
    store_texel(pf, px, c);
    size_t n = static_cast<size_t>(w) * h * samples, es = esize();
    for (size_t i = 0; i < n; ++i) std::memcpy(data.data() + i * es, px, es);
  }
};

struct texture_obj : object {
  int pf, samples;
  int max_lod = 0, min_lod = 0;
  uint32_t size[3];
  std::vector<ref<surface_obj>> surfs;
  texture_obj(int w, int h, int s, int pf_) : object(K_TEXTURE), pf(pf_), samples(s) {
    size[0] = w;
    size[1] = h;
    size[2] = 1;
    surfs.emplace_back(new surface_obj(w, h, s, pf_));
    release(surfs.back().get());  // ref took its own count
  }
  texture_view view() const {
    texture_view t;
    for (auto& s : surfs) t.levels.push_back(s->view());
    t.max_lod = max_lod;
    t.min_lod = min_lod;
    t.size[0] = size[0];
    t.size[1] = size[1];
    t.size[2] = size[2];
    return t;
  }
};

// surface::make_mip_surface (surface.cpp:53-92).  Odd sizes read one texel past the edge in the
// reference; the oracle clamps that read to the last texel.
surface_obj* make_mip(surface_obj* src, int filter) {
  int mw = (src->w + 1) / 2, mh = (src->h + 1) / 2;
  auto* r = new surface_obj(mw, mh, src->samples, src->pf);
  for (int y = 0; y < mh; ++y)
    for (int x = 0; x < mw; ++x)
      for (int s = 0; s < src->samples; ++s) {
        int x0 = x * 2, y0 = y * 2;
        int x1 = std::min(x0 + 1, src->w - 1), y1 = std::min(y0 + 1, src->h - 1);
        if (filter == FILTER_POINT) {
          r->set(x, y, s, src->get(x0, y0, s));
        } else {
          vec4 c0 = src->get(x0, y0, s), c1 = src->get(x1, y0, s), c2 = src->get(x0, y1, s), c3 = src->get(x1, y1, s);
          vec4 c = ((c0 + c1) + c2) + c3;
          r->set(x, y, s, c * 0.25f);
        }
      }
  return r;
}

struct sampler_obj : object {
  srx_sampler_desc desc;
  ref<texture_obj> tex;
  sampler_obj() : object(K_SAMPLER) {}
  sampler_params params() const {
    sampler_params p;
    p.min_filter = desc.min_filter;
    p.mag_filter = desc.mag_filter;
    p.mip_filter = desc.mip_filter;
    p.mip_qual = desc.mip_qual;
    p.addr_u = desc.addr_mode_u;
    p.addr_v = desc.addr_mode_v;
    p.lod_bias = desc.mip_lod_bias;
    p.max_aniso = desc.max_anisotropy;
    p.border = v4(desc.border_color[0], desc.border_color[1], desc.border_color[2], desc.border_color[3]);
    return p;
  }
};

struct layout_elem {
  std::string sem;  // lower case
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
struct query_obj : object {
  uint32_t id;
  bool running = false, has_data = false;
  srx_pipeline_statistics stats{};
  srx_internal_statistics istats{};
  uint64_t prof[8] = {};
  query_obj() : object(K_QUERY) {}
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
  if (l == "normal") return "normal";
  if (l == "texcoord") return "texcoord";
  return l;
}

// ---------------------------------------------------------------- shaders (sample shaders)
constexpr int MAX_REGS = 6;  // position + 5 attributes
struct vso {
  vec4 r[MAX_REGS];
};

struct vs_binding {
  const char* sem;
  uint32_t index;
  int reg;
};
// Register map of each built-in vertex shader (bind_semantic calls of the cpp shaders).
// (the SASL shaders: their input structs in declaration order)
int vs_bindings(int id, vs_binding* out) {
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
  case SRX_VS_BOX:  // TextureAndBlending.cpp:70-71
    out[0] = {"position", 0, 0};
    out[1] = {"texcoord", 0, 1};
    return 2;
  }
  return -1;
}
int vs_output_count(int id) {
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
size_t vs_const_bytes(int id) {
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
size_t ps_const_bytes(int id) {
  switch (id) {
  case SRX_PS_LAMBERT_TEX:
  case SRX_PS_CONSTANT:
  case SRX_PS_SHADOW_PCF: return 16;
  case SRX_PS_CUP: return 48;
  case SRX_PS_SSM: return 52;
  }
  return 0;
}
// eflib::transform(out, v, m) (eflib/src/math.cpp:142-154): out[i] = dot_prod4(v, column i)
vec4 transform_rowvec(vec4 v, const float* m) {
  vec4 o;
  for (int i = 0; i < 4; ++i) o[i] = v[0] * m[0 * 4 + i] + v[1] * m[1 * 4 + i] + v[2] * m[2 * 4 + i] + v[3] * m[3 * 4 + i];
  return o;
}
// eflib::transform33 (math.cpp:184-197): dot_prod3 with the first three columns, w = 0
vec4 transform33_rowvec(vec4 v, const float* m) {
  vec4 o;
  for (int i = 0; i < 3; ++i) o[i] = v[0] * m[0 * 4 + i] + v[1] * m[1 * 4 + i] + v[2] * m[2 * 4 + i];
  o[3] = 0.0f;
  return o;
}
vec4 sub4(const float* c, const vec4& v) { return v4(c[0] - v[0], c[1] - v[1], c[2] - v[2], c[3] - v[3]); }

void run_vs(int id, const float* consts, const vec4* in, vso& out) {
  const float* k = consts;
  switch (id) {
  case SRX_VS_COLORIZED:  // ColorizedTriangle.cpp:44-52 (vs_main in SASL)
    out.r[0] = transform_rowvec(in[0], k);
    out.r[1] = in[1];
    out.r[2] = sub4(k + 16, in[0]);
    out.r[3] = sub4(k + 20, in[0]);
    out.r[4] = sub4(k + 24, in[0]);
    return;
  case SRX_VS_MESH: {  // ComplexMesh.cpp:202-211: consts wv, proj, light_pos, eye_pos
    vec4 es = transform_rowvec(in[0], k);
    out.r[0] = transform_rowvec(es, k + 16);
    out.r[1] = v4(k[32] - es[0], k[33] - es[1], k[34] - es[2], 1.0f);
    out.r[2] = v4(k[36] - es[0], k[37] - es[1], k[38] - es[2], 1.0f);
    out.r[3] = transform33_rowvec(in[1], k);
    return;
  }
  case SRX_VS_SSM_DRAW:  // resources/ssm/Draw.savs: consts cameraPos, cameraWvp, lightPos, lightWvp
    out.r[0] = transform_rowvec(in[0], k + 4);
    out.r[1] = in[2];
    out.r[2] = in[1];
    out.r[3] = sub4(k + 20, in[0]);
    out.r[4] = sub4(k, in[0]);
    out.r[5] = transform_rowvec(in[0], k + 24);
    return;
  case SRX_VS_CUP:  // StencilMirror.cpp:47-55: consts wvp, eyePos, lightPos
    out.r[0] = transform_rowvec(in[0], k);
    out.r[1] = in[2];
    out.r[2] = in[1];
    out.r[3] = sub4(k + 20, in[0]);
    out.r[4] = sub4(k + 16, in[0]);
    return;
  case SRX_VS_BOX:  // TextureAndBlending.cpp:76-81
    out.r[0] = transform_rowvec(in[0], k);
    out.r[1] = in[0];
    out.r[2] = in[1];
    return;
  }
  switch (id) {
  case SRX_VS_PASSTHROUGH_COLOR:
    out.r[0] = in[0];
    out.r[1] = in[1];
    break;
  case SRX_VS_WVP_COLOR:
    out.r[0] = transform_rowvec(in[0], consts);
    out.r[1] = in[1];
    break;
  case SRX_VS_WVP_NORMAL_UV:
    out.r[0] = transform_rowvec(in[0], consts);
    out.r[1] = in[1];
    out.r[2] = in[2];
    break;
  case SRX_VS_WVP_COLOR_LIGHT:  // light-space position for the shadow lookup (StandardShadowMap.cpp)
    out.r[0] = transform_rowvec(in[0], consts);
    out.r[1] = in[1];
    out.r[2] = transform_rowvec(in[0], consts + 16);
    break;
  }
}

// eflib normalize3 (eflib/src/math.cpp:17-24)
void normalize3(const float* v, float* o) {
  float len = std::sqrt(length_sqr3(v[0], v[1], v[2]));
  if (equal_eps(len, 0.0f)) len = 1.0f;
  float inv = 1.0f / len;
  o[0] = v[0] * inv;
  o[1] = v[1] * inv;
  o[2] = v[2] * inv;
}

struct ps_context {
  int id;
  const float* consts;
  texture_view tex[SRX_MAX_SAMPLER_SLOTS];
  sampler_params sp[SRX_MAX_SAMPLER_SLOTS];
  bool has_tex[SRX_MAX_SAMPLER_SLOTS];
  bool sasl = false;         // a SASL pixel shader shades instead of id
  srx::sasl_program prog{};  // with its uniforms applied
};

// pixel_shader_unit::update + execute (shader_unit.cpp:106-176): input i of each pixel is vertex
// output attribute prog.inputs[i].attr, the program runs the quad, colour target 0 takes COLOR(0)
// (zero when the shader has none).
void run_sasl_quad(const ps_context& c, const vso* q, vec4* color) {
  using namespace srx;
  float in[4][SASL_MAX_INPUTS][4] = {};
  float out[4][SASL_MAX_OUTPUTS][4] = {};
  for (int px = 0; px < 4; ++px)
    for (uint32_t i = 0; i < c.prog.ninputs; ++i)
      for (int k = 0; k < 4; ++k) in[px][i][k] = q[px].r[1 + c.prog.inputs[i].attr][k];
  sasl_run_quad(c.prog, in, out, [&](int slot, int, float u, float v, float dxu, float dxv, float dyu, float dyv, float* r) {
    if (!c.has_tex[slot]) {
      r[0] = r[1] = r[2] = r[3] = 0.0f;
      return;
    }
    const vec4 t = sample_2d_grad(c.tex[slot], c.sp[slot], u, v, dxu, dxv, dyu, dyv, 0.0f);
    for (int k = 0; k < 4; ++k) r[k] = t[k];
  });
  for (int px = 0; px < 4; ++px) {
    color[px] = v4(0, 0, 0, 0);
    for (uint32_t o = 0; o < c.prog.noutputs; ++o)
      if (c.prog.outputs[o].target == 0) color[px] = v4(out[px][o][0], out[px][o][1], out[px][o][2], out[px][o][3]);
  }
}

// 3x3 percentage-closer shadow lookup (the tap layout of StandardShadowMap.cpp:86-106; each tap is a
// tex2dlod at lod 0 and is lit when the light-space depth minus the bias does not exceed it).
vec4 shadow_pcf(const ps_context& c, const vso& v) {
  const vec4& a = v.r[1];
  const vec4& l = v.r[2];
  float lx = l[0] / l[3], ly = l[1] / l[3], lz = l[2] / l[3];
  float cu = (lx + 1.0f) * 0.5f, cv = 1.0f - (ly + 1.0f) * 0.5f;
  float o = c.consts[2];
  int count = 0;
  if (c.has_tex[0]) {
    for (int ty = -1; ty <= 1; ++ty)
      for (int tx = -1; tx <= 1; ++tx) {
        float u = cu + static_cast<float>(tx) * o, vv = cv + static_cast<float>(ty) * o;
        vec4 t = sample_impl(c.tex[0], c.sp[0], u, vv, 0.0f, nullptr);
        if (lz - c.consts[1] <= t[0]) ++count;
      }
  }
  float lit = static_cast<float>(count) / 9.0f;
  float k = c.consts[0] + (1.0f - c.consts[0]) * lit;
  return v4(a[0] * k, a[1] * k, a[2] * k, a[3]);
}

// cpp_pixel_shader::tex2d(s, reg) (cpp_pixel_shader.cpp:13-31) on the four pixels of a quad: one LOD
// from the register's p1 - p0 and p2 - p0 (calc_lod_2d), then sample_2d_lod; no sampler bound:
// white (the shaders that test for one) .
void tex2d_cpp(const ps_context& c, int slot, const vso* q, int reg, vec4* out) {
  if (!c.has_tex[slot]) {
    for (int i = 0; i < 4; ++i) out[i] = v4(1, 1, 1, 1);
    return;
  }
  float lod = calc_lod_2d(c.tex[slot], c.sp[slot], q[1].r[reg][0] - q[0].r[reg][0], q[1].r[reg][1] - q[0].r[reg][1],
                          q[2].r[reg][0] - q[0].r[reg][0], q[2].r[reg][1] - q[0].r[reg][1]);
  for (int i = 0; i < 4; ++i) out[i] = sample_impl(c.tex[slot], c.sp[slot], q[i].r[reg][0], q[i].r[reg][1], lod, nullptr);
}
// eflib vector helpers on .xyz: dot_prod3, length, normalize3, reflect3 (math.cpp:17-41, 85-87;
// vector_generic.h:110-118).  Square roots are correctly rounded in float and in double alike.
inline float dot3(const vec4& a, const vec4& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline float length3(const vec4& a) {
  float t = 0.0f;
  t += a[0] * a[0];
  t += a[1] * a[1];
  t += a[2] * a[2];
  return std::sqrt(t);
}
inline vec4 normalize3v(const vec4& v) {
  float o[3];
  normalize3(v.v, o);
  return v4(o[0], o[1], o[2], 0.0f);
}
inline vec4 reflect3(const vec4& i, const vec4& n) {
  float k = 2.0f * dot3(i, n);
  return v4(i[0] - n[0] * k, i[1] - n[1] * k, i[2] - n[2] * k, 0.0f);
}

// ColorizedTriangle.cpp:55-92: attr0 normal, attr1..3 light vectors; 1/d^2 falloff, clamped
vec4 ps_colorized(const vso& v) {
  const vec4& n = v.r[1];
  float inv[3];
  vec4 nl[3];
  for (int i = 0; i < 3; ++i) {
    const vec4& l = v.r[2 + i];
    inv[i] = 1.0f / length3(l);
    nl[i] = v4(l[0] * inv[i], l[1] * inv[i], l[2] * inv[i], 0.0f);
  }
  vec4 nn = normalize3v(n);
  float refl[3];
  for (int i = 0; i < 3; ++i) refl[i] = dot3(nn, nl[i]);
  static const float col[3][3] = {{0.7f, 0.1f, 0.3f}, {0.1f, 0.3f, 0.7f}, {0.3f, 0.7f, 0.1f}};
  vec4 o = v4(0, 0, 0, 1.0f);
  for (int k = 0; k < 3; ++k) {
    float a = ((col[0][k] * refl[0]) * inv[0]) * inv[0];
    float b = ((col[1][k] * refl[1]) * inv[1]) * inv[1];
    float d = ((col[2][k] * refl[2]) * inv[2]) * inv[2];
    o[k] = std_clamp((a + b) + d, 0.0f, 1.0f);
  }
  return o;
}
// ComplexMesh.cpp:227-246 (std::pow(float, float))
vec4 ps_mesh(const vso& v) {
  vec4 l = normalize3v(v.r[1]), e = normalize3v(v.r[2]), n = normalize3v(v.r[3]);
  float n_dot_l = dot3(n, l);
  float roughness = 5;
  vec4 h = normalize3v(v4(l[0] + e[0], l[1] + e[1], l[2] + e[2], 0.0f));
  float spec = (roughness + 2) / 2 * std::pow(std_max(dot3(h, n), 0.0f), roughness);
  float clr = n_dot_l * (0.8f + spec * 0.4f);
  return v4(clr, clr, clr, 1.0f);
}
// StandardShadowMap.cpp:62-142: exponential shadow map over 3x3 tex2dlod taps (std::exp and std::log
// on float), white without a texture, pow(float, int) evaluated in double (C++11 <cmath>)
vec4 ps_ssm(const ps_context& c, const vso& v, const vec4& tex) {
  static const float esm_constant = 25000.0f;
  static const float gaussian_weights[9] = {0.027681f, 0.111014f, 0.027681f, 0.111014f, 0.445213f,
                                            0.111014f, 0.027681f, 0.111014f, 0.027681f};
  float occlusion = 0.0f;
  if (c.has_tex[0]) {
    const vec4& l4 = v.r[5];
    float lis[3] = {l4[0] / l4[3], l4[1] / l4[3], l4[2] / l4[3]};
    float cx = (lis[0] + 1.0f) * 0.5f, cy = (1.0f - (lis[1] + 1.0f) * 0.5f);
    float o = 1 / 512.0f;
    const float ox[9] = {-o, 0.0f, +o, -o, 0.0f, +o, -o, 0.0f, +o};
    const float oy[9] = {-o, -o, -o, 0.0f, 0.0f, 0.0f, +o, +o, +o};
    float sd[9];
    for (int i = 0; i < 9; ++i) sd[i] = sample_impl(c.tex[0], c.sp[0], cx + ox[i], cy + oy[i], 0.0f, nullptr)[0];
    float occluder = 0.0f;
    for (int i = 1; i < 9; ++i) occluder += gaussian_weights[i] * std::exp(esm_constant * (sd[i] - sd[0]));
    occluder += gaussian_weights[0];
    occluder = std::log(occluder);
    occluder += esm_constant * sd[0];
    occlusion = std_clamp(std::exp(occluder - esm_constant * lis[2]), 0.0f, 1.0f);
  }
  vec4 n = normalize3v(v.r[2]), ld = normalize3v(v.r[3]), ed = normalize3v(v.r[4]);
  float idiff = std_clamp(dot3(ld, n), 0.0f, 1.0f);
  vec4 r = reflect3(ld, n);
  float ispec = std_clamp(dot3(v4(-r[0], -r[1], -r[2], 0.0f), ed), 0.0f, 1.0f);
  int32_t shininess;
  std::memcpy(&shininess, &c.consts[12], 4);
  float p = static_cast<float>(std::pow(static_cast<double>(ispec), shininess));
  vec4 out;
  for (int k = 0; k < 3; ++k)
    out[k] = tex[k] * (c.consts[k] + (c.consts[4 + k] * idiff + c.consts[8 + k] * p) * occlusion);
  out[3] = 1.0f;
  return out;
}
// StencilMirror.cpp:66-110
vec4 ps_cup(const ps_context& c, const vso& v, const vec4& tex) {
  vec4 n = normalize3v(v.r[2]), ld = normalize3v(v.r[3]), ed = normalize3v(v.r[4]);
  float idiff = std_clamp(dot3(ld, n), 0.0f, 1.0f);
  float ispec = std_clamp(dot3(reflect3(ld, n), ed), 0.0f, 1.0f);
  vec4 out;
  for (int k = 0; k < 3; ++k)
    out[k] = tex[k] * ((c.consts[k] + c.consts[4 + k] * idiff) + c.consts[8 + k] * ispec);
  out[3] = 1.0f;
  return out;
}

// ddx/ddy of one value over a 2x2 quad (pixels 0 1 / 2 3) in a SASL pixel shader
// (sasl/src/codegen/cgs_simd.cpp:267-312): ddx pairs the columns of each row (p1 - p0 for pixels
// 0 and 1, p3 - p2 for 2 and 3), ddy the rows of each column (p2 - p0 for 0 and 2, p3 - p1 for 1
// and 3).  The cpp shaders use p1 - p0 and p2 - p0 for the whole quad instead.
void sasl_quad_derivatives(const float* p, float* ddx, float* ddy) {
  ddx[0] = ddx[1] = p[1] - p[0];
  ddx[2] = ddx[3] = p[3] - p[2];
  ddy[0] = ddy[2] = p[2] - p[0];
  ddy[1] = ddy[3] = p[3] - p[1];
}

// Runs the pixel shader on the 4 pixels of a quad (cpp_pixel_shader::execute,
// cpp_pixel_shader.cpp:62-74).  Returns the per-pixel pass mask (the sample shaders never discard).
void run_ps_quad(const ps_context& c, const vso* q, vec4* color) {
  if (c.sasl) return run_sasl_quad(c, q, color);
  switch (c.id) {
  case SRX_PS_GOURAUD:
    for (int i = 0; i < 4; ++i) color[i] = q[i].r[1];
    break;
  case SRX_PS_CONSTANT:
    for (int i = 0; i < 4; ++i) color[i] = v4(c.consts[0], c.consts[1], c.consts[2], c.consts[3]);
    break;
  case SRX_PS_LAMBERT_TEX: {
    // tex2D in a SASL pixel shader: the quad's derivatives (sasl_quad_derivatives), then
    // sample_2d_grad (sampler_api.cpp:13-18).
    float u[4], v[4], dxu[4], dyu[4], dxv[4], dyv[4];
    for (int i = 0; i < 4; ++i) {
      u[i] = q[i].r[2][0];
      v[i] = q[i].r[2][1];
    }
    sasl_quad_derivatives(u, dxu, dyu);
    sasl_quad_derivatives(v, dxv, dyv);
    for (int i = 0; i < 4; ++i) {
      vec4 t = c.has_tex[0] ? sample_2d_grad(c.tex[0], c.sp[0], u[i], v[i], dxu[i], dxv[i], dyu[i], dyv[i], 0.0f)
                            : v4(0, 0, 0, 0);
      float n[3];
      normalize3(q[i].r[1].v, n);
      float d = n[0] * c.consts[0] + n[1] * c.consts[1] + n[2] * c.consts[2];
      float k = std_clamp(d, 0.0f, 1.0f);
      color[i] = v4(t[0] * k, t[1] * k, t[2] * k, 1.0f);
    }
    break;
  }
  case SRX_PS_TEX_CPP: {
    // cpp tex2d: one LOD per quad from ddx = p1 - p0, ddy = p2 - p0 (cpp_pixel_shader.cpp:13-31)
    float lod = 0.0f;
    if (c.has_tex[0])
      lod = calc_lod_2d(c.tex[0], c.sp[0], q[1].r[1][0] - q[0].r[1][0], q[1].r[1][1] - q[0].r[1][1],
                        q[2].r[1][0] - q[0].r[1][0], q[2].r[1][1] - q[0].r[1][1]);
    for (int i = 0; i < 4; ++i) {
      vec4 t = c.has_tex[0] ? sample_impl(c.tex[0], c.sp[0], q[i].r[1][0], q[i].r[1][1], lod, nullptr) : v4(0, 0, 0, 0);
      t[3] = 1.0f;
      color[i] = t;
    }
    break;
  }
  case SRX_PS_COLORIZED:
    for (int i = 0; i < 4; ++i) color[i] = ps_colorized(q[i]);
    break;
  case SRX_PS_MESH:
    for (int i = 0; i < 4; ++i) color[i] = ps_mesh(q[i]);
    break;
  case SRX_PS_SSM: {
    vec4 tex[4];
    tex2d_cpp(c, 1, q, 1, tex);
    for (int i = 0; i < 4; ++i) color[i] = ps_ssm(c, q[i], tex[i]);
    break;
  }
  case SRX_PS_CUP: {
    vec4 tex[4];
    tex2d_cpp(c, 0, q, 1, tex);
    for (int i = 0; i < 4; ++i) color[i] = ps_cup(c, q[i], tex[i]);
    break;
  }
  case SRX_PS_BOX: {  // TextureAndBlending.cpp:96-114: tex2d(s0, attr1), alpha 0.5
    tex2d_cpp(c, 0, q, 2, color);
    for (int i = 0; i < 4; ++i) color[i][3] = 0.5f;
    break;
  }
  case SRX_PS_SHADOW_PCF:
    for (int i = 0; i < 4; ++i) color[i] = shadow_pcf(c, q[i]);
    break;
  default:
    for (int i = 0; i < 4; ++i) color[i] = v4(0, 0, 0, 0);
  }
}

// ---------------------------------------------------------------- renderer state
struct vb_slot {
  ref<buffer_obj> buf;
  size_t stride = 0, offset = 0;
};

// A compiled SASL pixel shader (the compiler is the product's front end; its programs are run by
// sasl_run_quad here).
struct shader_obj : object {
  srx::sasl_module mod;
  shader_obj() : object(K_SHADER) {}
};

struct render_state {
  vb_slot vbs[SRX_MAX_INPUT_SLOTS];
  ref<buffer_obj> ib;
  uint32_t index_fmt = SRX_FMT_R16_UINT;
  ref<layout_obj> layout;
  uint32_t topo = SRX_TOPOLOGY_TRIANGLE_LIST;
  int vs_id = 0, ps_id = 0, blend_id = SRX_BLEND_OFF;
  std::vector<float> vs_consts, ps_consts;
  ref<sampler_obj> samplers[SRX_MAX_SAMPLER_SLOTS];
  ref<sampler_obj> vs_samplers[SRX_MAX_SAMPLER_SLOTS];
  srx_raster_desc raster;
  srx_depth_stencil_desc ds;
  ref<raster_obj> raster_state;  // renderer_impl.cpp:276-277 creates the defaults
  ref<ds_obj> ds_state;
  int32_t stencil_ref = 0;
  std::vector<ref<surface_obj>> color_targets;
  ref<surface_obj> ds_target;
  srx_viewport target_vp{0, 0, 0, 0, 0, 0};
  size_t target_samples = 0;
  srx_viewport vp{0, 0, 0, 0, 0, 1};
  // set_pixel_shader_code (renderer_impl.cpp:310-316) and the constant buffer's variables and
  // samplers by name (shader_cbuffer.cpp, applied per draw by render_core.cpp:80-93)
  ref<shader_obj> ps_code;
  std::vector<std::pair<std::string, std::vector<uint8_t>>> ps_vars;
  std::vector<std::pair<std::string, ref<sampler_obj>>> ps_named_samplers;
};

struct tri_info {
  const vso* v0;
  bool front_face;
  float bbox[4];
  vec4 edge[3];
  vso ddx, ddy;
};

struct stats_acc {
  std::atomic<uint64_t> ia_vertices{0}, ia_primitives{0}, vs_invocations{0}, cinvocations{0}, cprimitives{0},
      ps_invocations{0}, backend_input_pixels{0};
};

struct batch_obj;
struct renderer_obj : object {
  render_state st;
  std::vector<ref<query_obj>> active;
  batch_obj* recording = nullptr;  // srx_batch_begin .. srx_batch_end
  // resource_manager (resource_manager.cpp:7-113): the mapped resource and map_read's copy
  int32_t map_mode = 0;
  ref<buffer_obj> mapped_buf;
  ref<surface_obj> mapped_surf;
  std::vector<uint8_t> map_copy;
  renderer_obj() : object(K_RENDERER) {
    srx_default_raster_desc(&st.raster);
    srx_default_depth_stencil_desc(&st.ds);
    auto* rs = new raster_obj();
    rs->desc = st.raster;
    st.raster_state = ref<raster_obj>(rs);
    release(rs);
    auto* ds = new ds_obj();
    ds->desc = st.ds;
    st.ds_state = ref<ds_obj>(ds);
    release(ds);
  }
  ~renderer_obj() override;
};

// ---------------------------------------------------------------- pipeline
constexpr int TILE_SIZE = 64;
constexpr uint64_t SAMPLE_MASK = 0xFFFF;
constexpr int MSC = 16;  // MAX_SAMPLE_COUNT

struct draw_ctx {
  render_state* st;
  int nattr;
  uint32_t attr_mod[5];
  int samples;
  uint64_t full_mask, quad_full_mask;
  float sp[16][2];
  float aa_dummy;
  // framebuffer
  surface_obj* ct[8];
  int nct;
  surface_obj* ds;
  bool early_z;
  bool read_depth, write_depth, stencil;
  uint32_t stencil_ref, read_mask, write_mask;
  ps_context ps;
  stats_acc* stats;
};

bool cmp(int f, float a, float b) {
  switch (f) {
  case SRX_CMP_NEVER: return false;
  case SRX_CMP_LESS: return a < b;
  case SRX_CMP_EQUAL: return a == b;
  case SRX_CMP_LESS_EQUAL: return a <= b;
  case SRX_CMP_GREATER: return a > b;
  case SRX_CMP_NOT_EQUAL: return a != b;
  case SRX_CMP_GREATER_EQUAL: return a >= b;
  }
  return true;
}
bool cmpu(int f, uint32_t a, uint32_t b) {
  switch (f) {
  case SRX_CMP_NEVER: return false;
  case SRX_CMP_LESS: return a < b;
  case SRX_CMP_EQUAL: return a == b;
  case SRX_CMP_LESS_EQUAL: return a <= b;
  case SRX_CMP_GREATER: return a > b;
  case SRX_CMP_NOT_EQUAL: return a != b;
  case SRX_CMP_GREATER_EQUAL: return a >= b;
  }
  return true;
}
// framebuffer.cpp:136-166
uint32_t stencil_op(int op, uint32_t ref_, uint32_t cur) {
  switch (op) {
  case SRX_SOP_KEEP: return cur;
  case SRX_SOP_ZERO: return 0;
  case SRX_SOP_REPLACE: return ref_;
  case SRX_SOP_INCR_SAT: return std::min<uint32_t>(0xFF, cur + 1);
  case SRX_SOP_DECR_SAT: return std::max<uint32_t>(0, cur - 1);
  case SRX_SOP_INVERT: return ~cur;
  case SRX_SOP_INCR_WRAP: return (cur + 1) & 0xFF;
  case SRX_SOP_DECR_WRAP: return (cur - 1 + 256) & 0xFF;
  }
  return cur;
}

struct framebuffer {
  draw_ctx* c;
  // ds sample access (depth_stencil_accessor<rg32f>, framebuffer.cpp:38-86)
  float* ds_addr(size_t x, size_t y, size_t s) { return reinterpret_cast<float*>(c->ds->addr(x, y, s)); }
  void read_ds(size_t x, size_t y, size_t s, float& d, uint32_t& st) {
    d = 0.0f;
    st = 0;
    if (!c->ds) return;
    float* p = ds_addr(x, y, s);
    if (c->read_depth) d = p[0];
    if (c->stencil) {
      uint32_t u;
      std::memcpy(&u, &p[1], 4);
      st = u & c->read_mask;
    }
  }
  void write_ds(size_t x, size_t y, size_t s, float d, uint32_t st, uint32_t wmask) {
    if (!c->ds) return;
    float* p = ds_addr(x, y, s);
    if (c->write_depth) p[0] = d;
    if (c->stencil) {
      uint32_t u = st & wmask;
      std::memcpy(&p[1], &u, 4);
    }
  }
  void blend(size_t x, size_t y, size_t s, vec4 src) {
    surface_obj* t = c->nct > 0 ? c->ct[0] : nullptr;
    if (!t) return;
    switch (c->st->blend_id) {
    case SRX_BLEND_OFF: t->set(x, y, s, src); break;
    case SRX_BLEND_ON: t->set(x, y, s, lerp4(t->get(x, y, s), src, src[3])); break;
    default: break;
    }
  }
  bool depth_test(float a, float b) { return c->st->ds.depth_enable ? cmp(c->st->ds.depth_func, a, b) : true; }
  // render_sample (framebuffer.cpp:445-479).  front_face is always false for rasterised
  // triangles, so the back-face stencil functions apply.
  void render_sample(size_t x, size_t y, size_t s, vec4 color, float depth, bool front) {
    if (c->early_z) {
      blend(x, y, s, color);
      return;
    }
    float od;
    uint32_t os;
    read_ds(x, y, s, od, os);
    bool dp = c->ds ? depth_test(depth, od) : true;
    const srx_depth_stencil_op_desc& f = front ? c->st->ds.front_face : c->st->ds.back_face;
    bool sp = c->stencil ? cmpu(f.stencil_func, c->stencil_ref, os) : true;
    if (dp && sp) {
      uint32_t ns = c->stencil ? stencil_op(f.stencil_pass_op, c->stencil_ref, os) : 0;
      blend(x, y, s, color);
      write_ds(x, y, s, depth, ns, c->write_mask);
    }
  }
  void render_sample_quad(size_t x, size_t y, uint64_t mask, const vec4* color, const float* depth, bool front,
                          const float* aa) {
    for (int i = 0; i < 4; ++i) {
      size_t px = x + (i & 1), py = y + ((i & 2) >> 1);
      uint64_t m = mask & SAMPLE_MASK;
      mask >>= MSC;
      if (!m) continue;
      if (px >= static_cast<size_t>(c->st->target_vp.w) || py >= static_cast<size_t>(c->st->target_vp.h)) continue;
      if (c->samples == 1) {
        render_sample(px, py, 0, color[i], depth[i], front);
      } else {
        for (int s = 0; s < c->samples; ++s)
          if (m & (1ull << s)) render_sample(px, py, s, color[i], depth[i] + aa[s], front);
      }
    }
  }
  // early_z_test (framebuffer.cpp:522-614): depth written at test time
  uint64_t early_z_pixel(size_t x, size_t y, uint32_t pmask, float depth, const float* aa) {
    if (px_out(x, y)) return 0;
    if (c->samples == 1) {
      float od;
      uint32_t os;
      read_ds(x, y, 0, od, os);
      if (depth_test(depth, od)) {
        write_ds(x, y, 0, depth, 0, 0);
        return 1;
      }
      return 0;
    }
    uint64_t m = 0;
    for (int s = 0; s < c->samples; ++s) {
      if (!(pmask & (1u << s))) continue;
      float od;
      uint32_t os;
      read_ds(x, y, s, od, os);
      float nd = aa[s] + depth;
      bool p = depth_test(nd, od);
      m |= static_cast<uint64_t>(p ? 1 : 0) << s;
      if (p) write_ds(x, y, s, nd, 0, 0);
    }
    return m;
  }
  bool px_out(size_t x, size_t y) {
    return x >= static_cast<size_t>(c->st->target_vp.w) || y >= static_cast<size_t>(c->st->target_vp.h);
  }
  uint64_t early_z_quad(size_t x, size_t y, uint64_t qmask, const float* depth, const float* aa) {
    uint64_t r = 0;
    for (int i = 0; i < 4; ++i) {
      uint32_t pm = static_cast<uint32_t>((qmask >> (MSC * i)) & SAMPLE_MASK);
      if (!pm) continue;
      r |= early_z_pixel(x + (i & 1), y + ((i & 2) >> 1), pm, depth[i], aa) << (MSC * i);
    }
    return r;
  }
};

struct raster {
  draw_ctx* c;
  framebuffer fb;
  std::vector<tri_info>* tris;
  uint64_t ps_inv = 0, be_pix = 0;

  // step_2d_unproj_pos_quad + step_2d_unproj_attr_n_quad (shader.cpp:257-367)
  void interp_quad(const tri_info& t, float dx, float dy, vso* q, bool attrs) {
    const vso& v0 = *t.v0;
    for (int k = 0; k < 4; ++k) {
      float a = v0.r[0][k] + ((t.ddx.r[0][k] * dx) + (t.ddy.r[0][k] * dy));
      q[0].r[0][k] = a;
      q[1].r[0][k] = a + t.ddx.r[0][k];
      q[2].r[0][k] = a + t.ddy.r[0][k];
      q[3].r[0][k] = q[1].r[0][k] + t.ddy.r[0][k];
    }
    if (!attrs) return;
    float inv_w[4];
    for (int i = 0; i < 4; ++i) inv_w[i] = 1.0f / q[i].r[0][3];
    for (int a = 0; a < c->nattr; ++a) {
      int R = a + 1;
      bool noint = c->attr_mod[a] & SRX_AM_NOINTERPOLATION;
      bool nopersp = c->attr_mod[a] & SRX_AM_NOPERSPECTIVE;
      for (int k = 0; k < 4; ++k) {
        float i00, i01, i10, i11;
        if (noint) {
          i00 = i01 = i10 = i11 = v0.r[R][k];
        } else {
          i00 = v0.r[R][k] + ((t.ddx.r[R][k] * dx) + (t.ddy.r[R][k] * dy));
          i01 = i00 + t.ddx.r[R][k];
          i10 = i00 + t.ddy.r[R][k];
          i11 = i01 + t.ddy.r[R][k];
        }
        if (nopersp) {
          q[0].r[R][k] = i00; q[1].r[R][k] = i01; q[2].r[R][k] = i10; q[3].r[R][k] = i11;
        } else {
          q[0].r[R][k] = i00 * inv_w[0];
          q[1].r[R][k] = i01 * inv_w[1];
          q[2].r[R][k] = i10 * inv_w[2];
          q[3].r[R][k] = i11 * inv_w[3];
        }
      }
    }
  }

  // draw_full_quad / draw_quad (rasterizer.cpp:1245-1421); full = quad fully covered
  void draw_quad(uint32_t left, uint32_t top, uint64_t qmask, bool full, const tri_info& t, const float* aa) {
    float dx = 0.5f + static_cast<float>(left) - t.v0->r[0][0];
    float dy = 0.5f + static_cast<float>(top) - t.v0->r[0][1];
    vso q[4];
    interp_quad(t, dx, dy, q, false);
    float depth[4] = {q[0].r[0][2], q[1].r[0][2], q[2].r[0][2], q[3].r[0][2]};
    uint64_t tested = qmask;
    if (c->early_z) tested = fb.early_z_quad(left, top, qmask, depth, aa);
    if (tested == 0) return;
    interp_quad(t, dx, dy, q, true);
    ps_inv += 4;
    vec4 color[4];
    run_ps_quad(c->ps, q, color);
    // the sample pixel shaders never return false, so the mask is unchanged
    uint64_t gate = full ? tested : qmask;  // draw_quad gates on the untested mask (:1409)
    if (gate != 0) {
      be_pix += 4;
      fb.render_sample_quad(left, top, tested, color, depth, t.front_face, aa);
    }
  }

  void draw_full_tile(int l, int t0, int r, int b, const tri_info& t, const float* aa) {
    for (int top = t0; top < b; top += 2)
      for (int left = l; left < r; left += 2) draw_quad(left, top, c->quad_full_mask, true, t, aa);
  }

  // draw_partial_tile (rasterizer.cpp:298-439), SSE branch order
  void draw_partial_tile(int left, int top, const tri_info& t, const float* aa) {
    float lf = static_cast<float>(left), tf = static_cast<float>(top);
    float ev[3];
    for (int e = 0; e < 3; ++e) ev[e] = t.edge[e][2] - ((lf * t.edge[e][0]) + (tf * t.edge[e][1]));
    uint32_t pm[16] = {};
    for (int s = 0; s < c->samples; ++s) {
      for (int iy = 0; iy < 4; ++iy) {
        float my = c->sp[s][1] + static_cast<float>(iy);
        for (int ix = 0; ix < 4; ++ix) {
          float mx = c->sp[s][0] + static_cast<float>(ix);
          bool rej = false;
          for (int e = 0; e < 3; ++e) rej |= ((mx * t.edge[e][0]) + (my * t.edge[e][1])) < ev[e];
          if (!rej) pm[iy * 4 + ix] |= 1u << s;
        }
      }
    }
    for (int quad = 0; quad < 4; ++quad) {
      int qx = (quad & 1) << 1, qy = quad & 2;
      int qs = qx | (qy << 2);
      uint64_t m = (static_cast<uint64_t>(pm[qs + 0]) & SAMPLE_MASK) | ((static_cast<uint64_t>(pm[qs + 1]) & SAMPLE_MASK) << 16) |
                   ((static_cast<uint64_t>(pm[qs + 4]) & SAMPLE_MASK) << 32) | ((static_cast<uint64_t>(pm[qs + 5]) & SAMPLE_MASK) << 48);
      if (m == 0) continue;
      if (m == c->quad_full_mask)
        draw_quad(left + qx, top + qy, c->quad_full_mask, true, t, aa);
      else
        draw_quad(left + qx, top + qy, m, false, t, aa);
    }
  }

  // subdivide_tile (rasterizer.cpp:441-602), SSE branch order
  void subdivide(int left, int top, uint32_t rx, uint32_t ry, uint32_t nw, uint32_t nh, const tri_info& t,
                 std::vector<uint32_t>& out, float xmin, float xmax, float ymin, float ymax, const float* rej2acc,
                 const float* evalue, const float* stepx, const float* stepy) {
    float lf = static_cast<float>(left), tf = static_cast<float>(top);
    float ev[3];
    for (int e = 0; e < 3; ++e) ev[e] = evalue[e] - ((lf * t.edge[e][0]) + (tf * t.edge[e][1]));
    for (int iy = 0; iy < 4; ++iy) {
      for (int ix = 0; ix < 4; ++ix) {
        bool rej = false, notacc = false;
        for (int e = 0; e < 3; ++e) {
          float sx = stepx[e] * static_cast<float>(ix);
          float sy = stepy[e] * static_cast<float>(iy);
          float r = sx + sy;
          float a = r + rej2acc[e];
          rej |= r < ev[e];
          notacc |= a < ev[e];
        }
        int32_t mx = static_cast<int32_t>(ix * nw + rx);
        int32_t my = static_cast<int32_t>(iy * nh + ry);
        rej |= xmin >= static_cast<float>(mx + static_cast<int32_t>(nw));
        rej |= xmax < static_cast<float>(mx);
        rej |= ymin >= static_cast<float>(my + static_cast<int32_t>(nh));
        rej |= ymax < static_cast<float>(my);
        if (!rej) out.push_back(static_cast<uint32_t>(mx) | (static_cast<uint32_t>(my) << 8) | (notacc ? 0u : 0x80000000u));
      }
    }
  }

  // rasterize_triangle (rasterizer.cpp:617-773)
  void rasterize_triangle(uint32_t rec, float tvx, float tvy) {
    uint32_t prim = rec >> 1, full = rec & 1;
    const tri_info& t = (*tris)[prim];
    bool mark_x[3], mark_y[3];
    for (int e = 0; e < 3; ++e) {
      mark_x[e] = t.edge[e][0] > 0;
      mark_y[e] = t.edge[e][1] > 0;
    }
    float xmin = t.bbox[0] - tvx, xmax = t.bbox[1] - tvx, ymin = t.bbox[2] - tvy, ymax = t.bbox[3] - tvy;
    std::vector<uint32_t> regions[2];
    regions[0].push_back(full << 31);
    int src = 0;
    int vpleft0 = fast_floori(tvx), vptop0 = fast_floori(tvy);
    uint32_t sw = static_cast<uint32_t>(fast_floori(static_cast<float>(TILE_SIZE)));
    uint32_t sh = sw;
    float stepx[3], stepy[3], rej2acc[3], evalue[3], part[3];
    for (int e = 0; e < 3; ++e) {
      stepx[e] = TILE_SIZE * t.edge[e][0];
      stepy[e] = TILE_SIZE * t.edge[e][1];
      rej2acc[e] = -std::fabs(stepx[e]) - std::fabs(stepy[e]);
      part[e] = static_cast<float>(mark_x[e] * TILE_SIZE) * t.edge[e][0] + static_cast<float>(mark_y[e] * TILE_SIZE) * t.edge[e][1];
      evalue[e] = t.edge[e][2] - part[e];
    }
    float aa[16];
    if (c->samples > 1) {
      for (int s = 0; s < c->samples; ++s)
        aa[s] = (c->sp[s][0] - 0.5f) * t.ddx.r[0][2] + (c->sp[s][1] - 0.5f) * t.ddy.r[0][2];
    } else {
      aa[0] = 0.0f;
    }
    float right = c->st->target_vp.w + c->st->target_vp.x;
    float bottom = c->st->target_vp.h + c->st->target_vp.y;
    while (!regions[src].empty()) {
      int dst = src ^ 1;
      regions[dst].clear();
      sw /= 4;
      sh /= 4;
      for (int e = 0; e < 3; ++e) {
        stepx[e] *= 0.25f;
        stepy[e] *= 0.25f;
        rej2acc[e] *= 0.25f;
        part[e] *= 0.25f;
        evalue[e] = t.edge[e][2] - part[e];
      }
      for (uint32_t packed : regions[src]) {
        uint32_t rx = packed & 0xFF, ry = (packed >> 8) & 0xFF;
        bool is_full = (packed >> 31) != 0;
        int vpleft = static_cast<int>(std::max(0u, static_cast<unsigned>(vpleft0 + static_cast<int>(rx))));
        int vptop = static_cast<int>(std::max(0u, static_cast<unsigned>(vptop0 + static_cast<int>(ry))));
        if (static_cast<float>(vpleft) >= right || static_cast<float>(vptop) >= bottom) continue;
        if (is_full) {
          int r = static_cast<int>(std::min(static_cast<uint32_t>(vpleft0) + rx + sw * 4, static_cast<uint32_t>(right)));
          int b = static_cast<int>(std::min(static_cast<uint32_t>(vptop0) + ry + sh * 4, static_cast<uint32_t>(bottom)));
          draw_full_tile(vpleft, vptop, r, b, t, aa);
        } else if (sw <= 1 && sh <= 1) {
          draw_partial_tile(vpleft, vptop, t, aa);
        } else {
          subdivide(vpleft, vptop, rx, ry, sw, sh, t, regions[dst], xmin, xmax, ymin, ymax, rej2acc, evalue, stepx, stepy);
        }
      }
      src = dst;
    }
  }
};

// compute_triangle_info (rasterizer.cpp:864-945)
void compute_triangle_info(tri_info& ti, const vso* const* verts, int nattr) {
  ti.v0 = nullptr;
  const vec4* p[3] = {&verts[0]->r[0], &verts[1]->r[0], &verts[2]->r[0]};
  // distance |x|+|y| evaluated in float, compared after widening to double
  double dist[3];
  for (int i = 0; i < 3; ++i) dist[i] = static_cast<double>(std::fabs((*p[i])[0]) + std::fabs((*p[i])[1]));
  int r0;
  if (dist[0] < dist[1]) r0 = dist[0] < dist[2] ? 0 : 2;
  else r0 = dist[1] < dist[2] ? 1 : 2;
  int r1 = (r0 + 1) % 3, r2 = (r1 + 1) % 3;
  const vso* rv[3] = {verts[r0], verts[r1], verts[r2]};
  vso e01, e02;
  for (int R = 0; R <= nattr; ++R) {
    e01.r[R] = rv[1]->r[R] - rv[0]->r[R];
    e02.r[R] = rv[2]->r[R] - rv[0]->r[R];
  }
  float area = e02.r[0][0] * e01.r[0][1] - e02.r[0][1] * e01.r[0][0];
  if (equal_eps(area, 0.0f)) return;
  ti.front_face = area > 0.0f;
  float inv_area = 1.0f / area;
  ti.bbox[0] = std_min(std_min((*p[0])[0], (*p[1])[0]), (*p[2])[0]);
  ti.bbox[1] = std_max(std_max((*p[0])[0], (*p[1])[0]), (*p[2])[0]);
  ti.bbox[2] = std_min(std_min((*p[0])[1], (*p[1])[1]), (*p[2])[1]);
  ti.bbox[3] = std_max(std_max((*p[0])[1], (*p[1])[1]), (*p[2])[1]);
  for (int i = 0; i < 3; ++i) {
    int s = i, e = (i + 1) % 3;
    ti.edge[i][0] = (*p[s])[1] - (*p[e])[1];
    ti.edge[i][1] = (*p[e])[0] - (*p[s])[0];
    ti.edge[i][2] = (*p[e])[0] * (*p[s])[1] - (*p[e])[1] * (*p[s])[0];
    ti.edge[i][3] = 0.0f;
  }
  // compute_derivative_n (shader.cpp:414-450)
  float e01x = e01.r[0][0], e01y = e01.r[0][1], e02x = e02.r[0][0], e02y = e02.r[0][1];
  for (int R = 0; R <= nattr; ++R)
    for (int k = 0; k < 4; ++k) {
      float xd = e02.r[R][k] * e01y - e01.r[R][k] * e02y;
      float yd = e01.r[R][k] * e02x - e02.r[R][k] * e01x;
      ti.ddx.r[R][k] = xd * inv_area;
      ti.ddy.r[R][k] = yd * inv_area;
    }
  ti.v0 = rv[0];
}

// lerp_n (shader.cpp:171-180)
void lerp_vso(vso& o, const vso& a, const vso& b, float t, int nattr, const uint32_t* mod) {
  o.r[0] = a.r[0] + (b.r[0] - a.r[0]) * t;
  for (int i = 0; i < nattr; ++i) {
    o.r[i + 1] = a.r[i + 1];
    if (!(mod[i] & SRX_AM_NOINTERPOLATION)) o.r[i + 1] = o.r[i + 1] + (b.r[i + 1] - a.r[i + 1]) * t;
  }
}

float dot4(vec4 a, vec4 b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3]; }

// raster_state cull functions (raster_state.cpp:10-32)
bool cull_fn(const srx_raster_desc& d, float area) {
  switch (d.cull_mode) {
  case SRX_CULL_NONE: return false;
  case SRX_CULL_FRONT: return d.front_ccw ? (area <= 0) : (area >= 0);
  case SRX_CULL_BACK: return d.front_ccw ? (area >= 0) : (area <= 0);
  }
  return false;
}

// clipper (clipper.cpp:43-228).  out: up to 9 vertex pointers; returns the count.
int clip_triangle(vso** tv, vso** out, std::deque<vso>& pool_, const srx_raster_desc& rd, int nattr, const uint32_t* mod) {
  const vec4 planes[2] = {v4(0, 0, 1, 0), v4(0, 0, -1, 1)};
  bool in = true;
  for (int pl = 0; pl < 2 && in; ++pl)
    for (int v = 0; v < 3; ++v)
      if (dot4(planes[pl], tv[v]->r[0]) < 0) {
        in = false;
        break;
      }
  if (in) {
    // compute_front (clipper.cpp:92-101)
    float pv[3][2];
    for (int v = 0; v < 3; ++v) {
      float iw = 1.0f / tv[v]->r[0][3];
      pv[v][0] = tv[v]->r[0][0] * iw;
      pv[v][1] = tv[v]->r[0][1] * iw;
    }
    float ax = pv[2][0] - pv[0][0], ay = pv[2][1] - pv[0][1], bx = pv[1][0] - pv[0][0], by = pv[1][1] - pv[0][1];
    bool front = (ax * by - ay * bx) > 0.0f;
    if (cull_fn(rd, front ? 1.0f : -1.0f)) return 0;
    int off = front ? 0 : 1;
    out[0] = tv[0];
    out[1] = tv[1 + off];
    out[2] = tv[2 - off];
    return 3;
  }
  vso* cv[2][5];
  uint32_t nc[2] = {3, 0};
  cv[0][0] = tv[0];
  cv[0][1] = tv[1];
  cv[0][2] = tv[2];
  int src = 0, dst = 1;
  bool front = false;
  for (int pl = 0; pl < 2; ++pl) {
    nc[dst] = 0;
    float d0 = 0.0f;
    if (nc[src] != 0) d0 = dot4(planes[pl], cv[src][0]->r[0]);
    for (uint32_t i = 0, j = 1; i < nc[src]; ++i, ++j) {
      j %= nc[src];
      float d1 = dot4(planes[pl], cv[src][j]->r[0]);
      if (d0 >= 0.0f) {
        cv[dst][nc[dst]++] = cv[src][i];
        if (d1 < 0.0f) {
          pool_.emplace_back();
          vso* pc = &pool_.back();
          lerp_vso(*pc, *cv[src][i], *cv[src][j], d0 / (d0 - d1), nattr, mod);
          cv[dst][nc[dst]++] = pc;
        }
      } else if (d1 >= 0.0f) {
        pool_.emplace_back();
        vso* pc = &pool_.back();
        lerp_vso(*pc, *cv[src][j], *cv[src][i], d1 / (d1 - d0), nattr, mod);
        cv[dst][nc[dst]++] = pc;
      }
      d0 = d1;
    }
    if (pl == 0 && nc[dst] >= 3) {
      float pv[3][2];
      for (int v = 0; v < 3; ++v) {
        float iaw = 1 / std::fabs(cv[dst][v]->r[0][3]);
        pv[v][0] = cv[dst][v]->r[0][0] * iaw;
        pv[v][1] = cv[dst][v]->r[0][1] * iaw;
      }
      float ax = pv[2][0] - pv[0][0], ay = pv[2][1] - pv[0][1], bx = pv[1][0] - pv[0][0], by = pv[1][1] - pv[0][1];
      float area = ax * by - ay * bx;
      front = area > 0.0f;
      if (cull_fn(rd, area)) return 0;
    }
    src ^= 1;
    dst ^= 1;
  }
  uint32_t n = nc[src];
  if (n < 3) return 0;
  // fan; the reference loops past n-2 triangles (clipper.cpp:79) -- only the counted ones are kept
  int cnt = 0;
  for (uint32_t t = 1; t + 1 < n; ++t) {
    out[cnt + 0] = cv[src][0];
    if (front) {
      out[cnt + 1] = cv[src][t];
      out[cnt + 2] = cv[src][t + 1];
    } else {
      out[cnt + 1] = cv[src][t + 1];
      out[cnt + 2] = cv[src][t];
    }
    cnt += 3;
  }
  return cnt;
}

// viewport_transform + project_n (shader.cpp:499-511, 117-135)
void viewport_project(vso& v, const srx_viewport& vp, int nattr, const uint32_t* mod) {
  vec4& p = v.r[0];
  float invw = equal_eps(p[3], 0.0f) ? 1.0f : 1.0f / p[3];
  vec4 pos = p * invw;
  float ox = (vp.x + vp.w) * 0.5f, oy = (vp.y + vp.h) * 0.5f;
  p[0] = (vp.w * 0.5f) * pos[0] + ox;
  p[1] = (vp.h * 0.5f) * -pos[1] + oy;
  p[2] = (vp.maxz - vp.minz) * pos[2] + vp.minz;
  p[3] = invw;
  for (int i = 0; i < nattr; ++i)
    if (!(mod[i] & SRX_AM_NOPERSPECTIVE)) v.r[i + 1] = v.r[i + 1] * p[3];
}

vec4 fetch_elem(const uint8_t* base, size_t size, size_t off, uint32_t fmt, float default_w) {
  float f[4] = {0, 0, 0, 0};
  size_t need = 0;
  switch (fmt) {
  case SRX_FMT_R32_FLOAT: need = 4; break;
  case SRX_FMT_R32G32_FLOAT: need = 8; break;
  case SRX_FMT_R32G32B32_FLOAT: need = 12; break;
  case SRX_FMT_R32G32B32A32_FLOAT:
  case SRX_FMT_R32G32B32A32_SINT: need = 16; break;
  }
  if (off + need > size) return v4(0, 0, 0, 0);  // the reference reads out of bounds here
  std::memcpy(f, base + off, need);
  // get_vec4 (stream_assembler.cpp:26-44)
  switch (fmt) {
  case SRX_FMT_R32_FLOAT: return v4(f[0], 0, 0, default_w);
  case SRX_FMT_R32G32_FLOAT: return v4(f[0], f[1], 0, default_w);
  case SRX_FMT_R32G32B32_FLOAT: return v4(f[0], f[1], f[2], default_w);
  default: return v4(f[0], f[1], f[2], f[3]);
  }
}

// A swap chain without a window (swap_chain_impl.cpp): back buffer and presented surface.
struct swap_obj : object {
  ref<renderer_obj> r;
  ref<texture_obj> back, resolved;
  swap_obj() : object(K_SWAP_CHAIN) {}
};

int32_t execute_draw(renderer_obj* r, bool indexed, uint32_t start, uint32_t primcnt, int32_t basevert) {
  render_state& st = r->st;
  if (st.color_targets.empty() && !st.ds_target) return SRX_OK;
  if (st.topo != SRX_TOPOLOGY_TRIANGLE_LIST && st.topo != SRX_TOPOLOGY_TRIANGLE_STRIP)
    return fail(SRX_E_UNSUPPORTED, "only triangle lists and strips are rasterised (lines are unimplemented upstream)");
  if (st.raster.fill_mode != SRX_FILL_SOLID) return fail(SRX_E_UNSUPPORTED, "wireframe fill is unimplemented upstream");
  int nout = vs_output_count(st.vs_id);
  if (nout <= 0) return fail(SRX_INVALID_PARAMETER, "no vertex shader bound");
  if (!st.layout) return fail(SRX_INVALID_PARAMETER, "no input layout");
  if (indexed && !st.ib) return fail(SRX_INVALID_PARAMETER, "no index buffer");
  if (st.ps_code)
    for (uint32_t i = 0; i < st.ps_code->mod.prog.ninputs; ++i)
      if (st.ps_code->mod.prog.inputs[i].attr >= nout)
        return fail(SRX_INVALID_PARAMETER, "the pixel shader reads more attributes than the vertex shader writes");

  // stream assembler register map (stream_assembler.cpp:57-86)
  vs_binding binds[8];
  int nb = vs_bindings(st.vs_id, binds);
  struct regsrc {
    const layout_elem* e;
    int reg;
    float dw;
  } rs[8];
  for (int i = 0; i < nb; ++i) {
    rs[i].e = nullptr;
    for (auto& e : st.layout->elems)
      if (e.sem == binds[i].sem && e.index == binds[i].index) rs[i].e = &e;
    if (!rs[i].e) return fail(SRX_INVALID_PARAMETER, "input layout lacks a semantic the vertex shader reads");
    rs[i].reg = binds[i].reg;
    rs[i].dw = std::string(binds[i].sem) == "position" ? 1.0f : 0.0f;
    if (!st.vbs[rs[i].e->slot].buf) return fail(SRX_INVALID_PARAMETER, "vertex buffer slot unbound");
  }

  draw_ctx c{};
  c.st = &st;
  c.nattr = nout;
  for (int i = 0; i < 5; ++i) c.attr_mod[i] = SRX_AM_LINEAR;
  c.samples = static_cast<int>(st.target_samples);
  // sample patterns (rasterizer.cpp:1086-1103)
  static const float p1[1][2] = {{0.5f, 0.5f}};
  static const float p2[2][2] = {{0.25f, 0.25f}, {0.75f, 0.75f}};
  static const float p4[4][2] = {{0.375f, 0.125f}, {0.875f, 0.375f}, {0.125f, 0.625f}, {0.625f, 0.875f}};
  const float(*pat)[2] = c.samples == 1 ? p1 : c.samples == 2 ? p2 : p4;
  if (c.samples != 1 && c.samples != 2 && c.samples != 4) return fail(SRX_E_UNSUPPORTED, "sample count must be 1, 2 or 4");
  for (int s = 0; s < c.samples; ++s) {
    c.sp[s][0] = pat[s][0];
    c.sp[s][1] = pat[s][1];
  }
  c.full_mask = (1ull << c.samples) - 1;
  c.quad_full_mask = c.full_mask | (c.full_mask << 16) | (c.full_mask << 32) | (c.full_mask << 48);
  c.nct = 0;
  for (auto& t : st.color_targets) c.ct[c.nct++] = t.get();
  c.ds = st.ds_target.get();
  const srx_depth_stencil_desc& ds = st.ds;
  c.stencil = c.ds && ds.stencil_enable;
  c.read_depth = c.ds && ds.depth_enable && ds.depth_func != SRX_CMP_NEVER && ds.depth_func != SRX_CMP_ALWAYS;
  c.write_depth = c.ds && ds.depth_enable && ds.depth_write_mask && ds.depth_func != SRX_CMP_NEVER;
  // early-Z is on iff stencil is off and the pixel shader writes no depth (framebuffer.cpp:424);
  // recomputed for every draw (the reference recomputes it only when the DS state changes).
  c.early_z = c.ds && !ds.stencil_enable;
  c.read_mask = ds.stencil_read_mask & 0xFF;
  c.write_mask = ds.stencil_write_mask & 0xFF;
  c.stencil_ref = c.stencil ? (static_cast<uint32_t>(st.stencil_ref) & c.read_mask) : 0;
  c.ps.id = st.ps_id;
  c.ps.consts = st.ps_consts.data();
  sampler_obj* bound[SRX_MAX_SAMPLER_SLOTS] = {};
  if (st.ps_code) {
    // SASL: its samplers are bound by name, slot k being its k-th sampler variable
    const srx::sasl_module& m = st.ps_code->mod;
    c.ps.sasl = true;
    c.ps.prog = m.prog;
    for (auto& v : st.ps_vars) srx::sasl_set_uniform(m, c.ps.prog, v.first, v.second.data(), v.second.size());
    for (size_t k = 0; k < m.samplers.size() && k < SRX_MAX_SAMPLER_SLOTS; ++k)
      for (auto& ns : st.ps_named_samplers)
        if (ns.first == m.samplers[k]) bound[k] = ns.second.get();
  } else {
    for (int i = 0; i < SRX_MAX_SAMPLER_SLOTS; ++i) bound[i] = st.samplers[i].get();
  }
  for (int i = 0; i < SRX_MAX_SAMPLER_SLOTS; ++i) {
    c.ps.has_tex[i] = bound[i] && bound[i]->tex;
    if (c.ps.has_tex[i]) {
      c.ps.tex[i] = bound[i]->tex->view();
      c.ps.sp[i] = bound[i]->params();
    }
  }

  stats_acc stats;
  c.stats = &stats;
  pool& pl = global_pool();
  unsigned nthreads = pl.size();

  // ---- index fetch + vertex shader through per-thread caches, clipping (geom_setup_engine.cpp)
  const uint8_t* ibuf = indexed ? st.ib->data.data() : nullptr;
  size_t ibuf_size = indexed ? st.ib->data.size() : 0;
  uint32_t istride = st.index_fmt == SRX_FMT_R16_UINT ? 2 : 4;
  size_t istart = static_cast<size_t>(start) * istride;
  auto fetch_index = [&](uint32_t prim, int k) -> uint32_t {
    uint32_t id;
    if (st.topo == SRX_TOPOLOGY_TRIANGLE_LIST) {
      id = prim * 3 + k;
    } else {
      uint32_t ids[3] = {prim, prim + 1, prim + 2};
      if (prim & 1) std::swap(ids[0], ids[2]);
      id = ids[k];
    }
    if (!indexed) return id + static_cast<uint32_t>(basevert);
    size_t off = istart + static_cast<size_t>(id) * istride;
    uint32_t v = 0;
    if (off + istride <= ibuf_size) {
      if (istride == 2) {
        uint16_t s16;
        std::memcpy(&s16, ibuf + off, 2);
        v = s16;
      } else {
        std::memcpy(&v, ibuf + off, 4);
      }
    }
    return v + static_cast<uint32_t>(basevert);
  };
  auto run_vertex = [&](uint32_t index, vso& out) {
    vec4 in[8];
    for (int i = 0; i < 8; ++i) in[i] = v4(0, 0, 0, 0);
    for (int i = 0; i < nb; ++i) {
      const vb_slot& vb = st.vbs[rs[i].e->slot];
      size_t off = rs[i].e->offset + vb.stride * index + vb.offset;
      in[rs[i].reg] = fetch_elem(vb.buf->data.data(), vb.buf->data.size(), off, rs[i].e->fmt, rs[i].dw);
    }
    for (int i = 0; i < MAX_REGS; ++i) out.r[i] = v4(0, 0, 0, 0);
    run_vs(st.vs_id, st.vs_consts.data(), in, out);
  };

  struct tcache {
    std::deque<vso> pool;
    std::pair<uint32_t, vso*> items[128];
    uint64_t vs_inv = 0;
  };
  std::vector<tcache> caches(nthreads);
  for (auto& tc : caches)
    for (auto& it : tc.items) it = {0xFFFFFFFFu, nullptr};
  size_t npkg = (primcnt + 7) / 8;
  std::vector<vso*> sparse(static_cast<size_t>(primcnt) * 9 + 9);
  std::vector<uint32_t> pkg_count(npkg, 0);
  std::atomic<uint64_t> cinv{0};
  pl.run(primcnt, 8, [&](unsigned tid, size_t b, size_t e) {
    tcache& tc = caches[tid];
    size_t pkg = b / 8;
    vso** cursor = sparse.data() + b * 9;
    uint32_t cnt = 0;
    for (size_t p = b; p < e; ++p) {
      vso* pv[3];
      for (int k = 0; k < 3; ++k) {
        uint32_t idx = fetch_index(static_cast<uint32_t>(p), k);
        auto& it = tc.items[idx % 128];
        if (it.first == idx) {
          pv[k] = it.second;
        } else {
          ++tc.vs_inv;
          tc.pool.emplace_back();
          vso* v = &tc.pool.back();
          run_vertex(idx, *v);
          it = {idx, v};
          pv[k] = v;
        }
      }
      int n = clip_triangle(pv, cursor, tc.pool, st.raster, nout, c.attr_mod);
      cursor += n;
      cnt += n;
    }
    pkg_count[pkg] = cnt;
    cinv.fetch_add(e - b);
  });
  // compaction (geom_setup_engine.cpp:126-171)
  std::vector<vso*> clipped;
  {
    std::vector<size_t> addr(npkg + 1, 0);
    for (size_t i = 0; i < npkg; ++i) addr[i + 1] = addr[i] + pkg_count[i];
    clipped.resize(addr[npkg]);
    for (size_t i = 0; i < npkg; ++i)
      std::memcpy(clipped.data() + addr[i], sparse.data() + i * 8 * 9, pkg_count[i] * sizeof(vso*));
  }
  size_t nclip = clipped.size() / 3;
  for (auto& tc : caches) stats.vs_invocations += tc.vs_inv;
  stats.ia_vertices += static_cast<uint64_t>(primcnt) * 3;
  stats.ia_primitives += primcnt;
  stats.cinvocations += cinv.load();
  stats.cprimitives += nclip;

  // ---- viewport transform of unique vertices (rasterizer.cpp:1203-1243)
  {
    std::vector<vso*> uniq(clipped);
    std::sort(uniq.begin(), uniq.end());
    uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
    pl.run(uniq.size(), 8, [&](unsigned, size_t b, size_t e) {
      for (size_t i = b; i < e; ++i) viewport_project(*uniq[i], st.vp, nout, c.attr_mod);
    });
  }

  // ---- triangle setup + 64x64 binning (rasterizer.cpp:775-862)
  size_t tx = static_cast<size_t>(st.vp.w + TILE_SIZE - 1) / TILE_SIZE;
  size_t ty = static_cast<size_t>(st.vp.h + TILE_SIZE - 1) / TILE_SIZE;
  size_t ntiles = tx * ty;
  std::vector<tri_info> tris(nclip);
  std::vector<std::vector<std::vector<uint32_t>>> bins(nthreads, std::vector<std::vector<uint32_t>>(ntiles));
  pl.run(nclip, 8, [&](unsigned tid, size_t b, size_t e) {
    auto& tb = bins[tid];
    for (size_t i = b; i < e; ++i) {
      const vso* v[3] = {clipped[i * 3], clipped[i * 3 + 1], clipped[i * 3 + 2]};
      tri_info& ti = tris[i];
      compute_triangle_info(ti, v, nout);
      if (!ti.v0) continue;
      int sx = std::min(fast_floori(std_max(0.0f, ti.bbox[0]) / TILE_SIZE), static_cast<int>(tx));
      int sy = std::min(fast_floori(std_max(0.0f, ti.bbox[2]) / TILE_SIZE), static_cast<int>(ty));
      int ex = std::min(fast_ceili(std_max(0.0f, ti.bbox[1]) / TILE_SIZE) + 1, static_cast<int>(tx));
      int ey = std::min(fast_ceili(std_max(0.0f, ti.bbox[3]) / TILE_SIZE) + 1, static_cast<int>(ty));
      uint32_t id = static_cast<uint32_t>(i);
      if (sx + 1 == ex && sy + 1 == ey) {
        tb[sy * tx + sx].push_back(id << 1);
        continue;
      }
      bool mx[3], my[3];
      float rej2acc[3];
      for (int k = 0; k < 3; ++k) {
        mx[k] = ti.edge[k][0] > 0;
        my[k] = ti.edge[k][1] > 0;
        float stx = TILE_SIZE * ti.edge[k][0], sty = TILE_SIZE * ti.edge[k][1];
        rej2acc[k] = -std::fabs(stx) - std::fabs(sty);
      }
      for (int y = sy; y < ey; ++y)
        for (int x = sx; x < ex; ++x) {
          bool rej = false, acc = true;
          for (int k = 0; k < 3; ++k) {
            float ev = ti.edge[k][2] - ((static_cast<float>(x + mx[k]) * TILE_SIZE) * ti.edge[k][0] +
                                        (static_cast<float>(y + my[k]) * TILE_SIZE) * ti.edge[k][1]);
            rej |= (0 < ev);
            acc &= (rej2acc[k] >= ev);
          }
          if (!rej) tb[y * tx + x].push_back((id << 1) | (acc ? 1u : 0u));
        }
    }
  });

  // ---- per-tile merge + sort + rasterize (rasterizer.cpp:947-1017)
  std::atomic<uint64_t> ps_inv{0}, be_pix{0};
  pl.run(ntiles, 1, [&](unsigned tid, size_t b, size_t e) {
    raster ras;
    ras.c = &c;
    ras.fb.c = &c;
    ras.tris = &tris;
    std::vector<uint32_t> prims;
    for (size_t tile = b; tile < e; ++tile) {
      prims.clear();
      for (unsigned t = 0; t < nthreads; ++t) prims.insert(prims.end(), bins[t][tile].begin(), bins[t][tile].end());
      std::sort(prims.begin(), prims.end());
      size_t yy = tile / tx, xx = tile - yy * tx;
      for (uint32_t rec : prims) ras.rasterize_triangle(rec, static_cast<float>(xx * TILE_SIZE), static_cast<float>(yy * TILE_SIZE));
    }
    ps_inv += ras.ps_inv;
    be_pix += ras.be_pix;
    (void)tid;
  });
  stats.ps_invocations += ps_inv.load();
  stats.backend_input_pixels += be_pix.load();

  for (auto& q : r->active) {
    if (q->id == SRX_QUERY_PIPELINE_STATISTICS) {
      q->stats.ia_vertices += stats.ia_vertices;
      q->stats.ia_primitives += stats.ia_primitives;
      q->stats.vs_invocations += stats.vs_invocations;
      q->stats.cinvocations += stats.cinvocations;
      q->stats.cprimitives += stats.cprimitives;
      q->stats.ps_invocations += stats.ps_invocations;
    } else if (q->id == SRX_QUERY_INTERNAL_STATISTICS) {
      q->istats.backend_input_pixels += stats.backend_input_pixels;
    }
  }
  return SRX_OK;
}

// ---------------------------------------------------------------- batches
// The batch contract is that every recorded command has the effect of its one-by-one execution;
// the restatement records each command with the state it saw and replays them in order.
struct batch_obj : object {
  renderer_obj* r = nullptr;
  std::vector<std::function<int32_t()>> cmds;
  uint32_t draws = 0, clears = 0, resolves = 0;
  uint64_t primitives = 0;
  batch_obj() : object(K_BATCH) {}
  ~batch_obj() override {
    if (r) release(r);
  }
};
renderer_obj::~renderer_obj() {
  if (recording) release(recording);
}

// The request server's contract is the plain pipeline applied to each request; the restatement
// renders every request at submission with the renderer API and reports it at once.
struct served {
  srx_request req;
  std::vector<uint8_t> color, depth;
  int64_t done_ns = 0;
  int32_t status = SRX_OK;
};
struct server_obj : object {
  srx_renderer r = nullptr;
  std::map<uint64_t, std::vector<float>> meshes;
  std::map<uint64_t, uint32_t> textures;
  std::map<uint64_t, served> finished;
  std::deque<uint64_t> unreported;
  uint64_t next_ticket = 1, jobs = 0, uploads = 0;
  std::chrono::steady_clock::time_point epoch = std::chrono::steady_clock::now();
  server_obj() : object(K_SERVER) {}
  ~server_obj() override { srx_release(r); }
  int64_t now() const {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - epoch).count();
  }
};

template <class T>
T* as(void* h, int kind) {
  auto* o = static_cast<object*>(h);
  return (o && o->kind == kind) ? static_cast<T*>(o) : nullptr;
}

}  // namespace

// ==================================================================== C ABI
extern "C" {

uint32_t srx_abi_version(void) { return SRX_ABI_VERSION; }
const char* srx_backend_name(void) { return "cpu"; }
// threads of the oracle's pool (reported beside the CPU baseline)
uint32_t srxo_thread_count(void) { return global_pool().size(); }
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
  *out = reinterpret_cast<srx_renderer>(static_cast<object*>(new renderer_obj()));
  return SRX_OK;
}

int32_t srx_create_buffer(srx_renderer r, size_t size, srx_buffer* out) {
  if (!as<renderer_obj>(r, K_RENDERER) || !out) return SRX_INVALID_PARAMETER;
  auto* b = new buffer_obj();
  b->data.assign(size, 0);
  *out = reinterpret_cast<srx_buffer>(static_cast<object*>(b));
  return SRX_OK;
}
size_t srx_buffer_size(srx_buffer b) {
  auto* o = as<buffer_obj>(b, K_BUFFER);
  return o ? o->data.size() : 0;
}
int32_t srx_buffer_write(srx_buffer b, size_t off, const void* data, size_t size) {
  auto* o = as<buffer_obj>(b, K_BUFFER);
  if (!o || off + size > o->data.size()) return SRX_INVALID_PARAMETER;
  std::memcpy(o->data.data() + off, data, size);
  return SRX_OK;
}
// buffer.cpp:42-62 (the contiguous case is the same row-by-row copy)
int32_t srx_buffer_transfer(srx_buffer b, size_t off, const void* data, size_t stride_dest, size_t stride_src, size_t size,
                            size_t count) {
  auto* o = as<buffer_obj>(b, K_BUFFER);
  if (!o || (!data && size && count)) return SRX_INVALID_PARAMETER;
  if (count == 0 || size == 0) return SRX_OK;
  if (off + stride_dest * (count - 1) + size > o->data.size()) return SRX_INVALID_PARAMETER;
  const uint8_t* src = static_cast<const uint8_t*>(data);
  for (size_t i = 0; i < count; ++i) std::memcpy(o->data.data() + off + i * stride_dest, src + i * stride_src, size);
  return SRX_OK;
}
int32_t srx_buffer_read(srx_buffer b, size_t off, void* data, size_t size) {
  auto* o = as<buffer_obj>(b, K_BUFFER);
  if (!o || off + size > o->data.size()) return SRX_INVALID_PARAMETER;
  std::memcpy(data, o->data.data() + off, size);
  return SRX_OK;
}

int32_t srx_create_tex2d(srx_renderer r, size_t w, size_t h, size_t ns, int32_t fmt, srx_texture* out) {
  if (!as<renderer_obj>(r, K_RENDERER) || !out || w == 0 || h == 0 || ns == 0 || fmt < 0 || fmt > 6)
    return SRX_INVALID_PARAMETER;
  *out = reinterpret_cast<srx_texture>(static_cast<object*>(new texture_obj(static_cast<int>(w), static_cast<int>(h),
                                                                            static_cast<int>(ns), fmt)));
  return SRX_OK;
}
int32_t srx_texture_surface(srx_texture t, size_t level, srx_surface* out) {
  auto* o = as<texture_obj>(t, K_TEXTURE);
  if (!o || !out || level >= o->surfs.size()) return SRX_INVALID_PARAMETER;
  // texture::subresource: only levels within [max_lod, min_lod]
  if (static_cast<int>(level) < o->max_lod || static_cast<int>(level) > o->min_lod) return SRX_INVALID_PARAMETER;
  retain(o->surfs[level].get());
  *out = reinterpret_cast<srx_surface>(static_cast<object*>(o->surfs[level].get()));
  return SRX_OK;
}
int32_t srx_texture_gen_mipmap(srx_texture t, int32_t filter, int32_t auto_gen) {
  auto* o = as<texture_obj>(t, K_TEXTURE);
  if (!o || (filter != SRX_FILTER_POINT && filter != SRX_FILTER_LINEAR)) return SRX_INVALID_PARAMETER;
  if (auto_gen) {
    // calc_lod_limit (texture.h:26-35)
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
    surface_obj* s = make_mip(o->surfs.back().get(), filter);
    o->surfs.emplace_back(s);
    release(s);
  }
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
size_t srx_surface_size_bytes(srx_surface s) { auto* o = as<surface_obj>(s, K_SURFACE); return o ? o->data.size() : 0; }
int32_t srx_surface_write(srx_surface s, const void* d, size_t n) {
  auto* o = as<surface_obj>(s, K_SURFACE);
  if (!o || n != o->data.size()) return SRX_INVALID_PARAMETER;
  std::memcpy(o->data.data(), d, n);
  return SRX_OK;
}
int32_t srx_surface_read(srx_surface s, void* d, size_t n) {
  auto* o = as<surface_obj>(s, K_SURFACE);
  if (!o || n != o->data.size()) return SRX_INVALID_PARAMETER;
  std::memcpy(d, o->data.data(), n);
  return SRX_OK;
}
// surface::resolve (surface.cpp:123-140)
int32_t srx_surface_resolve(srx_renderer r, srx_surface src, srx_surface dst) {
  auto* a = as<surface_obj>(src, K_SURFACE);
  auto* b = as<surface_obj>(dst, K_SURFACE);
  auto* ro = as<renderer_obj>(r, K_RENDERER);
  if (!ro || !a || !b || b->samples != 1 || a->w != b->w || a->h != b->h)
    return SRX_INVALID_PARAMETER;
  if (ro->recording) {
    batch_obj* bo = ro->recording;
    ref<surface_obj> ka(a), kb(b);
    bo->cmds.push_back([r, ka, kb]() {
      return srx_surface_resolve(r, reinterpret_cast<srx_surface>(static_cast<object*>(ka.get())),
                                 reinterpret_cast<srx_surface>(static_cast<object*>(kb.get())));
    });
    ++bo->resolves;
    return SRX_OK;
  }
  global_pool().run(static_cast<size_t>(a->h), 8, [&](unsigned, size_t y0, size_t y1) {
    for (size_t y = y0; y < y1; ++y)
      for (size_t x = 0; x < static_cast<size_t>(a->w); ++x) {
        vec4 c = v4(0, 0, 0, 0);
        for (int s = 0; s < a->samples; ++s) c = c + a->get(x, y, s);
        c = c * (1 / static_cast<float>(a->samples));
        b->set(x, y, 0, c);
      }
  });
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
  if (!r || start + n > SRX_MAX_INPUT_SLOTS) return SRX_INVALID_PARAMETER;
  for (size_t i = 0; i < n; ++i) {
    auto* b = as<buffer_obj>(bufs[i], K_BUFFER);
    r->st.vbs[start + i].buf = ref<buffer_obj>(b);
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
  if (size < vs_const_bytes(id) || (vs_const_bytes(id) && !consts)) return SRX_INVALID_PARAMETER;
  r->st.vs_id = id;
  r->st.vs_consts.assign(48, 0.0f);
  if (consts && size) std::memcpy(r->st.vs_consts.data(), consts, std::min<size_t>(size, 48 * sizeof(float)));
  return SRX_OK;
}
int32_t srx_set_pixel_shader(srx_renderer rr, int32_t id, const void* consts, size_t size) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r || id < SRX_PS_GOURAUD || id > SRX_PS_BOX) return SRX_INVALID_PARAMETER;
  if (size < ps_const_bytes(id) || (ps_const_bytes(id) && !consts)) return SRX_INVALID_PARAMETER;
  r->st.ps_id = id;
  r->st.ps_consts.assign(16, 0.0f);
  if (consts && size) std::memcpy(r->st.ps_consts.data(), consts, std::min<size_t>(size, 64));
  return SRX_OK;
}
int32_t srx_compile_shader(const char* code, size_t length, int32_t lang, srx_shader_code* out) {
  if (!code || !out) return SRX_INVALID_PARAMETER;
  *out = nullptr;
  if (lang != SRX_LANG_PIXEL_SHADER)
    return fail(SRX_E_UNSUPPORTED, "only pixel shaders are compiled; vertex shaders are the built-in ids");
  auto* o = new shader_obj();
  std::string msg;
  const int st = srx::sasl_compile_ps(code, length, o->mod, msg);
  if (st != srx::SASL_OK) {
    delete o;
    return fail(st == srx::SASL_UNSUPPORTED ? SRX_E_UNSUPPORTED : SRX_INVALID_PARAMETER, msg.c_str());
  }
  *out = reinterpret_cast<srx_shader_code>(static_cast<object*>(o));
  return SRX_OK;
}
int32_t srx_shader_reflect(srx_shader_code c, srx_shader_info* info) {
  auto* o = as<shader_obj>(c, K_SHADER);
  if (!o || !info) return SRX_INVALID_PARAMETER;
  std::memset(info, 0, sizeof *info);
  const srx::sasl_program& p = o->mod.prog;
  info->input_count = p.ninputs;
  info->output_count = p.noutputs;
  for (uint32_t i = 0; i < p.ninputs; ++i) info->input_components[i] = p.inputs[i].ncomp;
  for (uint32_t i = 0; i < p.noutputs; ++i) {
    info->output_targets[i] = p.outputs[i].target;
    info->output_components[i] = p.outputs[i].ncomp;
  }
  info->uniform_count = static_cast<uint32_t>(o->mod.uniforms.size());
  info->sampler_count = static_cast<uint32_t>(o->mod.samplers.size());
  info->instruction_count = p.ninst;
  info->register_count = p.nregs;
  return SRX_OK;
}
int32_t srx_set_pixel_shader_code(srx_renderer rr, srx_shader_code c) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r) return SRX_INVALID_PARAMETER;
  auto* o = c ? as<shader_obj>(c, K_SHADER) : nullptr;
  if (c && !o) return SRX_INVALID_PARAMETER;
  r->st.ps_code = ref<shader_obj>(o);
  return SRX_OK;
}
int32_t srx_get_pixel_shader_code(srx_renderer rr, srx_shader_code* out) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r || !out) return SRX_INVALID_PARAMETER;
  retain(r->st.ps_code.get());
  *out = reinterpret_cast<srx_shader_code>(static_cast<object*>(r->st.ps_code.get()));
  return SRX_OK;
}
int32_t srx_set_ps_variable(srx_renderer rr, const char* name, const void* data, size_t size) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r || !name || (size && !data)) return SRX_INVALID_PARAMETER;
  std::vector<uint8_t> bytes(static_cast<const uint8_t*>(data), static_cast<const uint8_t*>(data) + size);
  for (auto& v : r->st.ps_vars)
    if (v.first == name) {
      v.second = std::move(bytes);
      return SRX_OK;
    }
  r->st.ps_vars.emplace_back(name, std::move(bytes));
  return SRX_OK;
}
int32_t srx_set_ps_sampler_named(srx_renderer rr, const char* name, srx_sampler s) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r || !name) return SRX_INVALID_PARAMETER;
  auto* so = s ? as<sampler_obj>(s, K_SAMPLER) : nullptr;
  if (s && !so) return SRX_INVALID_PARAMETER;
  for (auto& v : r->st.ps_named_samplers)
    if (v.first == name) {
      v.second = ref<sampler_obj>(so);
      return SRX_OK;
    }
  r->st.ps_named_samplers.emplace_back(name, ref<sampler_obj>(so));
  return SRX_OK;
}
int32_t srx_ps_execute_quads(srx_renderer rr, srx_shader_code c, const float* in, float* out, size_t nquads) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  auto* o = as<shader_obj>(c, K_SHADER);
  if (!r || !o || (nquads && (!out || (!in && o->mod.prog.ninputs)))) return SRX_INVALID_PARAMETER;
  if (nquads > (1u << 24)) return fail(SRX_E_UNSUPPORTED, "more than 2^24 quads in one call");
  const srx::sasl_module& m = o->mod;
  ps_context pc{};
  pc.sasl = true;
  pc.prog = m.prog;
  for (auto& v : r->st.ps_vars) srx::sasl_set_uniform(m, pc.prog, v.first, v.second.data(), v.second.size());
  for (size_t k = 0; k < m.samplers.size() && k < SRX_MAX_SAMPLER_SLOTS; ++k)
    for (auto& ns : r->st.ps_named_samplers)
      if (ns.first == m.samplers[k] && ns.second && ns.second->tex) {
        pc.has_tex[k] = true;
        pc.tex[k] = ns.second->tex->view();
        pc.sp[k] = ns.second->params();
      }
  const uint32_t nin = pc.prog.ninputs, nout = pc.prog.noutputs;
  for (size_t qi = 0; qi < nquads; ++qi) {
    float qin[4][srx::SASL_MAX_INPUTS][4] = {};
    float qout[4][srx::SASL_MAX_OUTPUTS][4] = {};
    for (int px = 0; px < 4; ++px)
      for (uint32_t i = 0; i < nin; ++i)
        for (int k = 0; k < 4; ++k) qin[px][i][k] = in[((qi * 4 + px) * nin + i) * 4 + k];
    sasl_run_quad(pc.prog, qin, qout, [&](int slot, int, float u, float v, float dxu, float dxv, float dyu, float dyv, float* res) {
      if (!pc.has_tex[slot]) {
        res[0] = res[1] = res[2] = res[3] = 0.0f;
        return;
      }
      const vec4 t = sample_2d_grad(pc.tex[slot], pc.sp[slot], u, v, dxu, dxv, dyu, dyv, 0.0f);
      for (int k = 0; k < 4; ++k) res[k] = t[k];
    });
    for (int px = 0; px < 4; ++px)
      for (uint32_t oo = 0; oo < nout; ++oo)
        for (int k = 0; k < 4; ++k) out[((qi * 4 + px) * nout + oo) * 4 + k] = qout[px][oo][k];
  }
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
  r->st.raster_state = ref<raster_obj>(o);
  return SRX_OK;
}
int32_t srx_set_depth_stencil_state(srx_renderer rr, srx_ds_state d, int32_t ref_) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  auto* o = as<ds_obj>(d, K_DS);
  if (!r || !o) return SRX_INVALID_PARAMETER;
  r->st.ds = o->desc;
  r->st.ds_state = ref<ds_obj>(o);
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
// renderer_impl::set_viewport (renderer_impl.cpp:144-151)
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

int32_t srx_set_vs_sampler(srx_renderer rr, uint32_t slot, srx_sampler s) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r || slot >= SRX_MAX_SAMPLER_SLOTS) return SRX_INVALID_PARAMETER;
  r->st.vs_samplers[slot] = ref<sampler_obj>(as<sampler_obj>(s, K_SAMPLER));
  return SRX_OK;
}
// renderer_impl.cpp:61-157
int32_t srx_get_index_buffer(srx_renderer rr, srx_buffer* out) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r || !out) return SRX_INVALID_PARAMETER;
  retain(r->st.ib.get());
  *out = reinterpret_cast<srx_buffer>(static_cast<object*>(r->st.ib.get()));
  return SRX_OK;
}
uint32_t srx_get_index_format(srx_renderer rr) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  return r ? r->st.index_fmt : 0;
}
uint32_t srx_get_primitive_topology(srx_renderer rr) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  return r ? r->st.topo : 0;
}
int32_t srx_get_vertex_shader(srx_renderer rr) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  return r ? r->st.vs_id : 0;
}
int32_t srx_get_pixel_shader(srx_renderer rr) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  return r ? r->st.ps_id : 0;
}
int32_t srx_get_blend_shader(srx_renderer rr) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  return r ? r->st.blend_id : -1;
}
int32_t srx_get_rasterizer_state(srx_renderer rr, srx_raster_state* out) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r || !out) return SRX_INVALID_PARAMETER;
  retain(r->st.raster_state.get());
  *out = reinterpret_cast<srx_raster_state>(static_cast<object*>(r->st.raster_state.get()));
  return SRX_OK;
}
int32_t srx_get_depth_stencil_state(srx_renderer rr, srx_ds_state* out, int32_t* stencil_ref) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r || !out) return SRX_INVALID_PARAMETER;
  retain(r->st.ds_state.get());
  *out = reinterpret_cast<srx_ds_state>(static_cast<object*>(r->st.ds_state.get()));
  if (stencil_ref) *stencil_ref = r->st.stencil_ref;
  return SRX_OK;
}
int32_t srx_raster_state_desc(srx_raster_state rs, srx_raster_desc* d) {
  auto* o = as<raster_obj>(rs, K_RASTER);
  if (!o || !d) return SRX_INVALID_PARAMETER;
  *d = o->desc;
  return SRX_OK;
}
int32_t srx_depth_stencil_state_desc(srx_ds_state ds, srx_depth_stencil_desc* d) {
  auto* o = as<ds_obj>(ds, K_DS);
  if (!o || !d) return SRX_INVALID_PARAMETER;
  *d = o->desc;
  return SRX_OK;
}
// resource_manager::map_impl / buffer::map / surface::map (resource_manager.cpp:7-56,
// buffer.cpp:9-26, surface.cpp:94-116): map_read hands out a copy, the other modes the resource's
// own bytes; one mapping at a time.
static int32_t map_bytes(renderer_obj* r, srx_mapped_resource* m, std::vector<uint8_t>& bytes, int32_t mode,
                         uint32_t row_pitch, uint32_t depth_pitch) {
  if (mode < SRX_MAP_READ || mode > SRX_MAP_WRITE_NO_OVERWRITE) return SRX_FAILED;
  if (r->map_mode != 0) return SRX_FAILED;
  if (mode == SRX_MAP_READ) {
    r->map_copy = bytes;
    m->data = r->map_copy.data();
  } else {
    m->data = bytes.data();
  }
  m->row_pitch = row_pitch;
  m->depth_pitch = depth_pitch;
  r->map_mode = mode;
  return SRX_OK;
}
int32_t srx_map_buffer(srx_renderer rr, srx_mapped_resource* m, srx_buffer bb, int32_t mode) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  auto* b = as<buffer_obj>(bb, K_BUFFER);
  if (!r || !m || !b) return SRX_INVALID_PARAMETER;
  const uint32_t pitch = static_cast<uint32_t>(b->data.size());
  const int32_t rc = map_bytes(r, m, b->data, mode, pitch, pitch);
  if (rc == SRX_OK) r->mapped_buf = ref<buffer_obj>(b);
  return rc;
}
int32_t srx_map_surface(srx_renderer rr, srx_mapped_resource* m, srx_surface ss, int32_t mode) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  auto* o = as<surface_obj>(ss, K_SURFACE);
  if (!r || !m || !o) return SRX_INVALID_PARAMETER;
  const uint32_t pitch = static_cast<uint32_t>(static_cast<size_t>(o->w) * o->samples * o->esize());
  const int32_t rc = map_bytes(r, m, o->data, mode, pitch, pitch * static_cast<uint32_t>(o->h));
  if (rc == SRX_OK) r->mapped_surf = ref<surface_obj>(o);
  return rc;
}
int32_t srx_unmap(srx_renderer rr) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r) return SRX_INVALID_PARAMETER;
  if (r->map_mode == 0) return SRX_FAILED;
  r->map_mode = 0;
  r->mapped_buf = ref<buffer_obj>();
  r->mapped_surf = ref<surface_obj>();
  r->map_copy.clear();
  return SRX_OK;
}

int32_t srx_begin(srx_renderer rr, srx_query qq) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  auto* q = as<query_obj>(qq, K_QUERY);
  if (!r || !q) return SRX_INVALID_PARAMETER;
  if (q->running) return SRX_FAILED;
  for (auto& a : r->active)
    if (a->id == q->id) return SRX_FAILED;
  q->running = true;
  q->has_data = false;
  q->stats = srx_pipeline_statistics{};
  q->istats = srx_internal_statistics{};
  r->active.emplace_back(q);
  return SRX_OK;
}
int32_t srx_end(srx_renderer rr, srx_query qq) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  auto* q = as<query_obj>(qq, K_QUERY);
  if (!r || !q || !q->running) return SRX_FAILED;
  q->running = false;
  q->has_data = true;
  for (size_t i = 0; i < r->active.size(); ++i)
    if (r->active[i].get() == q) {
      r->active.erase(r->active.begin() + static_cast<long>(i));
      break;
    }
  return SRX_OK;
}
// The oracle executes every command when it is issued, so a finished query is always ready.
int32_t srx_get_data(srx_renderer rr, srx_query qq, void* data, size_t size, int32_t do_not_wait) {
  (void)do_not_wait;
  auto* q = as<query_obj>(qq, K_QUERY);
  if (!as<renderer_obj>(rr, K_RENDERER) || !q || !data) return SRX_ASYNC_ERROR;
  if (!q->has_data || q->running) return SRX_ASYNC_ERROR;
  if (q->id == SRX_QUERY_PIPELINE_STATISTICS && size >= sizeof(q->stats)) std::memcpy(data, &q->stats, sizeof(q->stats));
  else if (q->id == SRX_QUERY_INTERNAL_STATISTICS && size >= sizeof(q->istats)) std::memcpy(data, &q->istats, sizeof(q->istats));
  else if (q->id == SRX_QUERY_PIPELINE_PROFILES && size >= sizeof(q->prof)) std::memcpy(data, q->prof, sizeof(q->prof));
  else return SRX_ASYNC_ERROR;
  return SRX_ASYNC_READY;
}
static int32_t record_draw(renderer_obj* r, bool indexed, uint32_t start, uint32_t primcnt, int32_t basevert) {
  batch_obj* b = r->recording;
  render_state st = r->st;
  b->cmds.push_back([r, st, indexed, start, primcnt, basevert]() mutable {
    std::swap(r->st, st);
    const int32_t rc = execute_draw(r, indexed, start, primcnt, basevert);
    std::swap(r->st, st);
    return rc;
  });
  ++b->draws;
  b->primitives += primcnt;
  return SRX_OK;
}
int32_t srx_draw(srx_renderer rr, size_t start, size_t primcnt) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r) return SRX_INVALID_PARAMETER;
  if (r->recording) return record_draw(r, false, 0, static_cast<uint32_t>(primcnt), 0);
  // renderer_impl::draw: base_vertex 0; the index fetcher ignores start for non-indexed draws
  (void)start;
  return execute_draw(r, false, 0, static_cast<uint32_t>(primcnt), 0);
}
// The oracle has no device memory to avoid allocating: the same draw, refused inside a batch.
int32_t srx_draw_noalloc(srx_renderer rr, size_t start, size_t primcnt) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r) return SRX_INVALID_PARAMETER;
  if (r->recording) return SRX_E_UNSUPPORTED;
  (void)start;
  return execute_draw(r, false, 0, static_cast<uint32_t>(primcnt), 0);
}
int32_t srx_draw_index(srx_renderer rr, size_t start, size_t primcnt, int32_t basevert) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r) return SRX_INVALID_PARAMETER;
  if (r->recording) return record_draw(r, true, static_cast<uint32_t>(start), static_cast<uint32_t>(primcnt), basevert);
  return execute_draw(r, true, static_cast<uint32_t>(start), static_cast<uint32_t>(primcnt), basevert);
}
int32_t srx_clear_color(srx_renderer rr, srx_surface s, const float* c) {
  auto* o = as<surface_obj>(s, K_SURFACE);
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r || !o || !c) return SRX_INVALID_PARAMETER;
  if (r->recording) {
    ref<surface_obj> keep(o);
    const float cc[4] = {c[0], c[1], c[2], c[3]};
    r->recording->cmds.push_back([keep, cc]() {
      keep->fill(v4(cc[0], cc[1], cc[2], cc[3]));
      return SRX_OK;
    });
    ++r->recording->clears;
    return SRX_OK;
  }
  o->fill(v4(c[0], c[1], c[2], c[3]));
  return SRX_OK;
}
// render_core::clear_depth_stencil (render_core.cpp:101-111), framebuffer::clear_depth_stencil (:616-644)
int32_t srx_clear_depth_stencil(srx_renderer rr, srx_surface s, uint32_t flags, float d, uint32_t st) {
  auto* o = as<surface_obj>(s, K_SURFACE);
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r || !o || o->pf != PF_RG32F) return SRX_INVALID_PARAMETER;
  if (r->recording) {
    batch_obj* b = r->recording;
    r->recording = nullptr;  // validate now, replay later
    ref<surface_obj> keep(o);
    b->cmds.push_back([rr, keep, flags, d, st]() {
      return srx_clear_depth_stencil(rr, reinterpret_cast<srx_surface>(static_cast<object*>(keep.get())), flags, d, st);
    });
    ++b->clears;
    r->recording = b;
    return SRX_OK;
  }
  if (flags == (SRX_CLEAR_DEPTH | SRX_CLEAR_STENCIL)) {
    float sf;
    std::memcpy(&sf, &st, 4);
    o->fill(v4(d, sf, 0, 0));
    return SRX_OK;
  }
  if (flags != SRX_CLEAR_DEPTH && flags != SRX_CLEAR_STENCIL) return SRX_E_UNSUPPORTED;
  size_t n = static_cast<size_t>(o->w) * o->h * o->samples;
  float* p = reinterpret_cast<float*>(o->data.data());
  for (size_t i = 0; i < n; ++i) {
    if (flags == SRX_CLEAR_DEPTH) p[i * 2] = d;
    else std::memcpy(&p[i * 2 + 1], &st, 4);
  }
  return SRX_OK;
}
int32_t srx_flush(srx_renderer rr) { return as<renderer_obj>(rr, K_RENDERER) ? SRX_OK : SRX_INVALID_PARAMETER; }

int32_t srx_create_swap_chain(srx_renderer rr, size_t w, size_t h, size_t ns, int32_t fmt, srx_swap_chain* out) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r || !out) return SRX_INVALID_PARAMETER;
  srx_texture back = nullptr, res = nullptr;
  int32_t e = srx_create_tex2d(rr, w, h, ns, fmt, &back);
  if (e != SRX_OK) return e;
  if (ns > 1) {
    e = srx_create_tex2d(rr, w, h, 1, fmt, &res);
    if (e != SRX_OK) {
      srx_release(back);
      return e;
    }
  }
  auto* sc = new swap_obj();
  sc->r = ref<renderer_obj>(r);
  sc->back = ref<texture_obj>(as<texture_obj>(back, K_TEXTURE));
  sc->resolved = ref<texture_obj>(as<texture_obj>(res ? res : back, K_TEXTURE));
  srx_release(back);
  if (res) srx_release(res);
  *out = reinterpret_cast<srx_swap_chain>(static_cast<object*>(sc));
  return SRX_OK;
}
int32_t srx_swap_chain_surface(srx_swap_chain s, srx_surface* out) {
  auto* sc = as<swap_obj>(s, K_SWAP_CHAIN);
  if (!sc || !out) return SRX_INVALID_PARAMETER;
  return srx_texture_surface(reinterpret_cast<srx_texture>(static_cast<object*>(sc->back.get())), 0, out);
}
int32_t srx_swap_chain_presented(srx_swap_chain s, srx_surface* out) {
  auto* sc = as<swap_obj>(s, K_SWAP_CHAIN);
  if (!sc || !out) return SRX_INVALID_PARAMETER;
  return srx_texture_surface(reinterpret_cast<srx_texture>(static_cast<object*>(sc->resolved.get())), 0, out);
}
// swap_chain_impl::present (swap_chain_impl.cpp:41-49) without present_impl's window copy
int32_t srx_swap_chain_present(srx_swap_chain s) {
  auto* sc = as<swap_obj>(s, K_SWAP_CHAIN);
  if (!sc) return SRX_INVALID_PARAMETER;
  auto* rr = reinterpret_cast<srx_renderer>(static_cast<object*>(sc->r.get()));
  int32_t e = srx_flush(rr);
  if (e != SRX_OK || sc->resolved.get() == sc->back.get()) return e;
  srx_surface a = nullptr, b = nullptr;
  e = srx_swap_chain_surface(s, &a);
  if (e == SRX_OK) e = srx_swap_chain_presented(s, &b);
  if (e == SRX_OK) e = srx_surface_resolve(rr, a, b);
  if (a) srx_release(a);
  if (b) srx_release(b);
  return e;
}

void srx_default_server_desc(srx_server_desc* d) {
  *d = srx_server_desc{200ull << 30, 64ull << 30, 2ull << 30, 64, 3, 8192, 64};
}
int32_t srx_create_server(srx_renderer rr, const srx_server_desc* d, srx_server* out) {
  if (!as<renderer_obj>(rr, K_RENDERER) || !d || !out) return SRX_INVALID_PARAMETER;
  auto* S = new server_obj();
  srx_retain(rr);
  S->r = rr;
  *out = reinterpret_cast<srx_server>(static_cast<object*>(S));
  return SRX_OK;
}
int32_t srx_server_put_mesh(srx_server ss, uint64_t key, const void* vertices, uint32_t n_tris) {
  auto* S = as<server_obj>(ss, K_SERVER);
  if (!S || n_tris == 0) return SRX_INVALID_PARAMETER;
  auto it = S->meshes.find(key);
  if (it != S->meshes.end()) return it->second.size() == n_tris * 24ull ? SRX_OK : SRX_INVALID_PARAMETER;
  if (!vertices) return SRX_INVALID_PARAMETER;
  const float* f = static_cast<const float*>(vertices);
  S->meshes[key].assign(f, f + n_tris * 24ull);
  ++S->uploads;
  return SRX_OK;
}
int32_t srx_server_put_texture(srx_server ss, uint64_t key, const void* rgba8, uint32_t size) {
  auto* S = as<server_obj>(ss, K_SERVER);
  if (!S || key == 0 || size == 0) return SRX_INVALID_PARAMETER;
  if (!S->textures.count(key)) {
    if (!rgba8) return SRX_INVALID_PARAMETER;
    S->textures[key] = size;
    ++S->uploads;
  }
  return SRX_OK;
}
static int32_t serve_one(server_obj* S, const srx_request& q, served& o) {
  srx_renderer r = S->r;
  const std::vector<float>& m = S->meshes[q.mesh_key];
  srx_buffer vb = nullptr;
  srx_texture ct = nullptr, dt = nullptr, rt = nullptr;
  srx_surface cs = nullptr, ds = nullptr, rs = nullptr;
  srx_input_layout lay = nullptr;
  srx_raster_state ras = nullptr;
  srx_ds_state dss = nullptr;
  int32_t rc = srx_create_buffer(r, m.size() * 4, &vb);
  if (rc == SRX_OK) rc = srx_buffer_write(vb, 0, m.data(), m.size() * 4);
  if (rc == SRX_OK) rc = srx_create_tex2d(r, q.width, q.height, q.samples, SRX_PF_RGBA8, &ct);
  if (rc == SRX_OK) rc = srx_create_tex2d(r, q.width, q.height, q.samples, SRX_PF_RG32F, &dt);
  if (rc == SRX_OK) rc = srx_texture_surface(ct, 0, &cs);
  if (rc == SRX_OK) rc = srx_texture_surface(dt, 0, &ds);
  if (rc == SRX_OK && q.samples > 1) rc = srx_create_tex2d(r, q.width, q.height, 1, SRX_PF_RGBA8, &rt);
  if (rc == SRX_OK && q.samples > 1) rc = srx_texture_surface(rt, 0, &rs);
  srx_input_element_desc el[2] = {{"POSITION", 0, SRX_FMT_R32G32B32A32_FLOAT, 0, 0, 0, 0},
                                  {"COLOR", 0, SRX_FMT_R32G32B32A32_FLOAT, 0, 16, 0, 0}};
  if (rc == SRX_OK) rc = srx_create_input_layout(r, el, 2, SRX_VS_PASSTHROUGH_COLOR, &lay);
  srx_raster_desc rd;
  srx_default_raster_desc(&rd);
  rd.cull_mode = SRX_CULL_NONE;
  if (rc == SRX_OK) rc = srx_create_raster_state(r, &rd, &ras);
  srx_depth_stencil_desc dd;
  srx_default_depth_stencil_desc(&dd);
  if (rc == SRX_OK) rc = srx_create_depth_stencil_state(r, &dd, &dss);
  if (rc == SRX_OK) {
    size_t stride = 32, offs = 0;
    srx_set_input_layout(r, lay);
    srx_set_vertex_buffers(r, 0, 1, &vb, &stride, &offs);
    srx_set_primitive_topology(r, SRX_TOPOLOGY_TRIANGLE_LIST);
    srx_set_vertex_shader(r, SRX_VS_PASSTHROUGH_COLOR, nullptr, 0);
    srx_set_pixel_shader(r, SRX_PS_GOURAUD, nullptr, 0);
    srx_set_blend_shader(r, SRX_BLEND_OFF);
    srx_set_rasterizer_state(r, ras);
    srx_set_depth_stencil_state(r, dss, 0);
    srx_set_render_targets(r, 1, &cs, ds);
    srx_viewport vp{0, 0, static_cast<float>(q.width), static_cast<float>(q.height), 0, 1};
    srx_set_viewport(r, &vp);
    rc = srx_clear_color(r, cs, q.clear_color);
    if (rc == SRX_OK) rc = srx_clear_depth_stencil(r, ds, SRX_CLEAR_DEPTH | SRX_CLEAR_STENCIL, 1.0f, 0);
    if (rc == SRX_OK) rc = srx_draw(r, 0, q.n_tris);
    if (rc == SRX_OK && q.samples > 1) rc = srx_surface_resolve(r, cs, rs);
  }
  if (rc == SRX_OK) {
    o.color.resize(static_cast<size_t>(q.width) * q.height * 4);
    rc = srx_surface_read(q.samples > 1 ? rs : cs, o.color.data(), o.color.size());
  }
  if (rc == SRX_OK && q.want_depth) {
    o.depth.resize(static_cast<size_t>(q.width) * q.height * q.samples * 8);
    rc = srx_surface_read(ds, o.depth.data(), o.depth.size());
  }
  for (void* h : {static_cast<void*>(vb), static_cast<void*>(ct), static_cast<void*>(dt), static_cast<void*>(rt),
                  static_cast<void*>(cs), static_cast<void*>(ds), static_cast<void*>(rs), static_cast<void*>(lay),
                  static_cast<void*>(ras), static_cast<void*>(dss)})
    if (h) srx_release(h);
  return rc;
}
int32_t srx_server_submit(srx_server ss, const srx_request* q, uint64_t* ticket) {
  auto* S = as<server_obj>(ss, K_SERVER);
  if (!S || !q || !ticket || q->width == 0 || q->height == 0 || q->n_tris == 0 ||
      !(q->samples == 1 || q->samples == 2 || q->samples == 4))
    return SRX_INVALID_PARAMETER;
  int32_t rc = srx_server_put_mesh(ss, q->mesh_key, q->mesh_data, q->n_tris);
  if (rc == SRX_OK && q->texture_key) rc = srx_server_put_texture(ss, q->texture_key, q->texture_data, q->texture_size);
  if (rc != SRX_OK) return rc;
  served o;
  o.req = *q;
  o.status = serve_one(S, *q, o);
  o.done_ns = S->now();
  *ticket = S->next_ticket++;
  S->finished[*ticket] = std::move(o);
  S->unreported.push_back(*ticket);
  ++S->jobs;
  return SRX_OK;
}
int32_t srx_server_poll(srx_server ss, srx_completion* out, size_t capacity, size_t* count, int32_t wait_ms) {
  (void)wait_ms;
  auto* S = as<server_obj>(ss, K_SERVER);
  if (!S || !count || (capacity && !out)) return SRX_INVALID_PARAMETER;
  size_t n = 0;
  while (n < capacity && !S->unreported.empty()) {
    const uint64_t tk = S->unreported.front();
    S->unreported.pop_front();
    auto it = S->finished.find(tk);
    if (it == S->finished.end()) continue;  // released before it was reported
    const served& o = it->second;
    srx_completion& c = out[n++];
    c = srx_completion{};
    c.ticket = tk;
    c.status = o.status;
    c.batch_jobs = 1;
    c.arrival_ns = o.req.arrival_ns;
    c.admitted_ns = o.req.arrival_ns;
    c.done_ns = o.done_ns;
    c.color = o.color.empty() ? nullptr : o.color.data();
    c.color_bytes = o.color.size();
    c.depth = o.depth.empty() ? nullptr : o.depth.data();
    c.depth_bytes = o.depth.size();
  }
  *count = n;
  return SRX_OK;
}
int32_t srx_server_release(srx_server ss, uint64_t ticket) {
  auto* S = as<server_obj>(ss, K_SERVER);
  if (!S || !S->finished.erase(ticket)) return SRX_INVALID_PARAMETER;
  S->unreported.erase(std::remove(S->unreported.begin(), S->unreported.end(), ticket), S->unreported.end());
  return SRX_OK;
}
int64_t srx_server_now(srx_server ss) {
  auto* S = as<server_obj>(ss, K_SERVER);
  return S ? S->now() : -1;
}
int32_t srx_server_get_stats(srx_server ss, srx_server_stats* st) {
  auto* S = as<server_obj>(ss, K_SERVER);
  if (!S || !st) return SRX_INVALID_PARAMETER;
  *st = srx_server_stats{};
  st->batches = S->jobs;
  st->jobs = S->jobs;
  st->uploads = S->uploads;
  st->max_batch_jobs = S->jobs ? 1 : 0;
  st->resident_meshes = S->meshes.size();
  st->resident_textures = S->textures.size();
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
  retain(r);
  b->r = r;
  *out = reinterpret_cast<srx_batch>(static_cast<object*>(b));
  return SRX_OK;
}
int32_t srx_batch_launch(srx_batch bb) {
  auto* b = as<batch_obj>(bb, K_BATCH);
  if (!b) return SRX_INVALID_PARAMETER;
  batch_obj* rec = b->r->recording;
  b->r->recording = nullptr;
  int32_t rc = SRX_OK;
  for (auto& c : b->cmds)
    if ((rc = c()) != SRX_OK) break;
  b->r->recording = rec;
  return rc;
}
int32_t srx_batch_status(srx_batch bb, uint32_t* status) {
  if (!as<batch_obj>(bb, K_BATCH) || !status) return SRX_INVALID_PARAMETER;
  *status = 0;
  return SRX_OK;
}
int32_t srx_batch_configure(srx_batch bb, int32_t tile, int32_t bin_items, int32_t vis_threads) {
  (void)tile;
  (void)bin_items;
  (void)vis_threads;
  return as<batch_obj>(bb, K_BATCH) ? SRX_OK : SRX_INVALID_PARAMETER;
}
int32_t srx_batch_autotune(srx_batch bb, uint32_t reps, srx_batch_tuning* results, size_t capacity, size_t* count) {
  (void)reps;
  (void)results;
  (void)capacity;
  if (!as<batch_obj>(bb, K_BATCH)) return SRX_INVALID_PARAMETER;
  if (count) *count = 0;  // nothing to tune on the host
  return SRX_OK;
}
int32_t srx_batch_get_info(srx_batch bb, srx_batch_info* info) {
  auto* b = as<batch_obj>(bb, K_BATCH);
  if (!b || !info) return SRX_INVALID_PARAMETER;
  *info = srx_batch_info{};
  info->jobs = b->draws;
  info->clears = b->clears;
  info->resolves = b->resolves;
  info->primitives = b->primitives;
  return SRX_OK;
}

}  // extern "C"

// the sampler reads weights through this table
namespace srxo {
float ewa(int i) { return ewa_table[i]; }
}

// ---- test-only entry points (the oracle's helpers, for the pinning tests) ----
extern "C" {
int32_t srxo_test_fast_floori(double d) { return fast_floori(d); }
int32_t srxo_test_fast_ceili(double d) { return fast_ceili(d); }
float srxo_test_fast_log2(float v) { return fast_log2(v); }
float srxo_test_fast_round(float v) { return fast_round(v); }
float srxo_test_ewa(int32_t i) { return ewa_table[i & 255]; }
// Triangle setup of three screen-space positions (x, y, z, w each): edges[9], bbox[4], v0 index,
// returns 0 for a degenerate triangle.
// The SASL derivatives of a 4x4 pixel package (row-major, 4 per line) taken quad by quad.
void srxo_test_sasl_ddx_ddy(const float* in16, float* ddx16, float* ddy16) {
  for (int qy = 0; qy < 4; qy += 2)
    for (int qx = 0; qx < 4; qx += 2) {
      const int idx[4] = {qy * 4 + qx, qy * 4 + qx + 1, (qy + 1) * 4 + qx, (qy + 1) * 4 + qx + 1};
      float p[4], dx[4], dy[4];
      for (int i = 0; i < 4; ++i) p[i] = in16[idx[i]];
      sasl_quad_derivatives(p, dx, dy);
      for (int i = 0; i < 4; ++i) {
        ddx16[idx[i]] = dx[i];
        ddy16[idx[i]] = dy[i];
      }
    }
}
// The sampler of a sampler object at explicit coordinates (sampler.cpp:680-873):
// mode 0: sample_2d_grad(u, v, a=ddx.u, b=ddx.v, c=ddy.u, d=ddy.v, bias); mode 1: sample_impl at level
// a with no anisotropic information.
int32_t srxo_test_sample(srx_sampler ss, int32_t mode, float u, float v, float a, float b, float c, float d,
                         float bias, float* out4) {
  auto* so = as<sampler_obj>(ss, K_SAMPLER);
  if (!so || !so->tex || !out4) return SRX_INVALID_PARAMETER;
  const texture_view t = so->tex->view();
  const sampler_params p = so->params();
  const vec4 r = mode == 0 ? sample_2d_grad(t, p, u, v, a, b, c, d, bias) : sample_impl(t, p, u, v, a, nullptr);
  for (int k = 0; k < 4; ++k) out4[k] = r[k];
  return SRX_OK;
}
// calc_lod (sampler.cpp:521-603) and calc_anisotropic_info (:875-958) of a sampler object's texture:
// out = lod, then (anisotropic) probe count, weight_D, delta u, delta v.
int32_t srxo_test_lod(srx_sampler ss, int32_t aniso, float dxu, float dxv, float dyu, float dyv, float bias, float* out5) {
  auto* so = as<sampler_obj>(ss, K_SAMPLER);
  if (!so || !so->tex || !out5) return SRX_INVALID_PARAMETER;
  const texture_view t = so->tex->view();
  const sampler_params p = so->params();
  const vec4 ddx = v4(dxu, dxv, 0, 0), ddy = v4(dyu, dyv, 0, 0);
  if (aniso) {
    const aniso_info a = calc_aniso(t, p, ddx, ddy, bias);
    out5[0] = a.lod;
    out5[1] = a.probe_count;
    out5[2] = a.weight_D;
    out5[3] = a.delta_uv[0];
    out5[4] = a.delta_uv[1];
  } else {
    out5[0] = calc_lod(t, p, ddx, ddy, bias);
    out5[1] = out5[2] = out5[3] = out5[4] = 0.0f;
  }
  return SRX_OK;
}
int32_t srxo_test_triangle_setup(const float* pos12, float* edges9, float* bbox4, int32_t* v0_index) {
  vso v[3];
  for (int i = 0; i < 3; ++i) {
    for (int k = 0; k < 4; ++k) v[i].r[0][k] = pos12[i * 4 + k];
    for (int r = 1; r < MAX_REGS; ++r) v[i].r[r] = v4(0, 0, 0, 0);
  }
  const vso* vp[3] = {&v[0], &v[1], &v[2]};
  tri_info t;
  compute_triangle_info(t, vp, 0);
  if (!t.v0) return 0;
  for (int e = 0; e < 3; ++e)
    for (int k = 0; k < 3; ++k) edges9[e * 3 + k] = t.edge[e][k];
  for (int k = 0; k < 4; ++k) bbox4[k] = t.bbox[k];
  *v0_index = static_cast<int32_t>(t.v0 - &v[0]);
  return 1;
}
}
